/*
 * ctsm_gpu.h -- C-ABI of the MI355X column-biogeophysics library (libctsm_mi355x.so).
 *
 * The library replaces three routines of CTSM's per-time-step physics with
 * HIP kernels for gfx950, behind the same data the Fortran routines read and
 * write:
 *
 *   SoilWater       src/biogeophys/SoilWaterMovementMod.F90:240   (moisture form, :976-1440)
 *   SoilTemperature src/biogeophys/SoilTemperatureMod.F90:92
 *   CanopyFluxes    src/biogeophys/CanopyFluxesMod.F90:191
 *
 * Data conventions (mirror the reference's derived types, see
 * src/main/ColumnType.F90:137-155, src/biogeophys/SoilStateType.F90:134-161,
 * TemperatureType.F90:228-301):
 *
 *  - Every array is the Fortran array itself: the pointer is the address of
 *    its FIRST element as allocated by the reference's InitAllocate
 *    (c_loc(arr) on the Fortran side).  Column arrays are (begc:endc) or
 *    (begc:endc, lo:hi) with the column index fastest; element (c, j) sits at
 *    ptr[(c - begc_alloc) + (j - lo) * ldc].  Patch arrays likewise with
 *    begp_alloc / ldp.
 *  - The lower level bound `lo` of each 2-D field is fixed by this header
 *    (comment on each member): SNO = -nlevsno+1, ZI = -nlevsno, SOI = 1,
 *    SNOW_ONLY = -nlevsno+1 (upper bound 0), LYR = -nlevsno+1 (upper bound 1).
 *  - Column / patch numbers in filters and in patch->column, patchi, patchf
 *    are the reference's global 1-based indices (the ones stored in
 *    filter_nolakec etc.), not offsets.
 *  - All physics is IEEE binary64.  Integers are int32.
 *
 * Error behaviour: every entry point returns 0 on success or a CTSM_E_* code.
 * When a column fails inside a kernel (a singular pivot in the tridiagonal or
 * banded LU, which the reference turns into endrun at
 * SoilWaterMovementMod.F90:1295 / BandDiagonalMod.F90:212) the call returns
 * CTSM_E_SOLVER and ctsm_last_error() reports the first failing column so
 * the Fortran wrapper can call endrun(subgrid_index=c, ...) as the reference
 * does.
 *
 * Threading: a ctsm_ctx owns one HIP stream and its scratch; calls on
 * different contexts may run concurrently (one context per OpenMP clump,
 * as clm_driver.F90:525 runs clumps).  Calls on one context are serialised
 * by the caller.
 */
#ifndef CTSM_GPU_H
#define CTSM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------- */
#define CTSM_OK            0
#define CTSM_E_ARG         1   /* bad argument (null pointer, size, layout)      */
#define CTSM_E_CONFIG      2   /* namelist combination the library does not run  */
#define CTSM_E_HIP         3   /* HIP runtime error                              */
#define CTSM_E_SOLVER      4   /* LAPACK-style info != 0 in a column solve        */
#define CTSM_E_NODEVICE    5   /* no GPU visible                                  */
#define CTSM_E_ENDRUN      6   /* a check where the reference calls endrun fails   */
#define CTSM_E_IO          7   /* checkpoint file cannot be written / read, or does
                                  not hold the fields asked for (name, size, checksum) */
#define CTSM_E_SCRATCH     8   /* the call needs more than the fixed scratch of a
                                  *_noalloc entry (ctsm_last_error: the KiB needed) */
                               /* (canopy: quadratic, brent bracket, gs < 0, ...)  */

/* soil water lower / upper boundary conditions (SoilWaterMovementMod.F90:55-59) */
#define CTSM_BC_HEAD        0
#define CTSM_BC_FLUX        1
#define CTSM_BC_ZERO_FLUX   2
#define CTSM_BC_WATERTABLE  3
#define CTSM_BC_AQUIFER     4

/* snow thermal conductivity (SoilTemperatureMod.F90:739-760) */
#define CTSM_SNOWTHK_JORDAN1991 0
#define CTSM_SNOWTHK_STURM1997  1

/* CanopyFluxes / Photosynthesis switches (values as in the reference) */
#define CTSM_Z0_ZENGWANG2007      0   /* z0param_method = 'ZengWang2007' (CanopyFluxesMod.F90:905-911) */
#define CTSM_Z0_MEIER2022         1   /* z0param_method = 'Meier2022'    (CanopyFluxesMod.F90:913-932) */
#define CTSM_STOMATA_BB1987       1   /* stomatalcond_mtd_bb1987     (PhotosynthesisMod.F90:82) */
#define CTSM_STOMATA_MEDLYN2011   2   /* stomatalcond_mtd_medlyn2011 (PhotosynthesisMod.F90:83) */
#define CTSM_SOILRESIS_LEEPIELKE1992 0 /* soil_resis_method (SurfaceResistanceMod.F90:23-24)       */
#define CTSM_SOILRESIS_SL14       1

/* ---- namelist / parameter block -----------------------------------------
 * Replaces init_soilwater_movement (SoilWaterMovementMod.F90:110-216),
 * CanopyFluxesReadNML (CanopyFluxesMod.F90:92-156), the layer dimensions of
 * clm_varpar_init (clm_varpar.F90:290-304) and the param-file scalar e_ice
 * (SoilWaterMovementMod.F90:89-105).  Filled by ctsm_config_default(). */
typedef struct ctsm_config {
    int32_t nlevsno;            /* 12 (namelist_defaults_ctsm.xml:511)              */
    int32_t nlevsoi;            /* 20 (20SL_8.5m)                                   */
    int32_t nlevgrnd;           /* 25                                               */
    int32_t soilwater_movement_method; /* 1 = moisture form (only form supported) */
    int32_t upper_boundary_condition;  /* CTSM_BC_FLUX                            */
    int32_t lower_boundary_condition;  /* CTSM_BC_ZERO_FLUX, CTSM_BC_FLUX or
     * CTSM_BC_WATERTABLE (SoilWaterMovementMod.F90:1376-1377, 1767-1808 and
     * compute_qcharge :2016-2135): every column must then have nbedrock =
     * nlevsoi (the reference requires use_bedrock = .false., :181-183; a
     * column that does not fails with CTSM_E_ARG) and col->zwt / col->smpmin;
     * the water-table and drainage entries, which follow the path without the
     * aquifer layer, return CTSM_E_CONFIG on such a context               */
    int32_t flux_calculation;   /* 1 = inexpensive, 42 = expensive                  */
    int32_t snow_thermal_cond_method;  /* CTSM_SNOWTHK_*                          */
    int32_t itmax_canopy_fluxes;/* 40                                               */
    int32_t pad0;
    double  dtmin;              /* 60 s                                             */
    double  verySmall;          /* 1e-8                                             */
    double  xTolerUpper;        /* 0.1                                              */
    double  xTolerLower;        /* 0.01                                             */
    double  e_ice;              /* 6 (ice impedance exponent)                       */
    /* CanopyFluxes / Photosynthesis namelist switches.  use_fates, use_cn,
     * use_luna, use_c13, lnc_opt and perchroot are .false. and vcmax_opt = 0,
     * nlevcan = 1 on this path. */
    int32_t z0param_method;     /* CTSM_Z0_*                        (xml:621-622)    */
    int32_t stomatalcond_method;/* CTSM_STOMATA_*                   (xml:108)        */
    int32_t soil_resis_method;  /* CTSM_SOILRESIS_*                 (xml:270)        */
    int32_t use_biomass_heat_storage;  /* 0/1                       (xml:457)        */
    int32_t use_undercanopy_stability; /* 0/1                       (xml:454)        */
    int32_t light_inhibit;      /* 0/1                              (xml:95)         */
    /* use_hydrstress (clm_varctl, namelist_defaults_ctsm.xml:726): 1 selects
     * PhotosynthesisHydraulicStress and the PHS branches of CanopyFluxes
     * (CanopyFluxesMod.F90:638, 1133, 1157, 1219, 1353).  With 1, CanopyFluxes
     * runs through ctsm_canopyfluxes_phs_dev (the other canopy and land-step
     * entries return CTSM_E_CONFIG); with 0, ctsm_canopyfluxes_phs_dev does. */
    int32_t use_hydrstress;     /* 0                                                 */
    /* use_excess_ice (clm_varctl; SoilTemperatureMod.F90:287 and the excess-ice
     * terms of SoilThermProp / Phasechange): ctsm_create returns CTSM_E_CONFIG
     * for 1. */
    int32_t use_excess_ice;     /* 0                                                 */
    /* modifyphoto_and_lmr_forcrop (photosyns_inst, namelist_defaults_ctsm.xml:
     * 104): 0/1, read by PhotosynthesisHydraulicStress only
     * (PhotosynthesisMod.F90:3516, 3533, 3739, 3774). */
    int32_t modifyphoto_and_lmr_forcrop;   /* 1                                      */
    int32_t pad1;
} ctsm_config;

/* ---- array layout of one call -------------------------------------------- */
typedef struct ctsm_layout {
    int32_t begc_alloc;   /* lower column bound of the allocation (bounds_proc%begc) */
    int32_t ldc;          /* column extent of the allocation                          */
    int32_t begp_alloc;   /* lower patch bound of the allocation                      */
    int32_t ldp;          /* patch extent of the allocation                           */
    int32_t begg_alloc;   /* lower gridcell bound of the allocation (canopy path only) */
    int32_t ldg;          /* gridcell extent of the allocation                        */
} ctsm_layout;

/* ---- column state touched by the three routines --------------------------
 * Member comments: [lo] lower level bound, (in)/(out)/(inout), reference
 * owner type.  A member the called routine does not use may be NULL. */
typedef struct ctsm_col_state {
    /* ColumnType (col%) */
    const int32_t *snl;           /* (c)      in   number of snow layers (<= 0)          */
    const int32_t *nbedrock;      /* (c)      in   bedrock index = Richards layers        */
    const int32_t *patchi;        /* (c)      in   first patch of the column              */
    const int32_t *patchf;        /* (c)      in   last patch of the column               */
    const double  *dz;            /* [SNO]    in   layer thickness (m)                    */
    const double  *z;             /* [SNO]    in   node depth (m)                         */
    const double  *zi;            /* [ZI]     in   interface depth (m)                    */
    /* SoilStateType (soilstate_inst%) */
    const double  *watsat;        /* [SOI]    in   porosity                               */
    const double  *bsw;           /* [SOI]    in   Clapp-Hornberger b                     */
    const double  *sucsat;        /* [SOI]    in   saturated suction (mm)                 */
    const double  *hksat;         /* [SOI]    in   saturated conductivity (mm/s)          */
    const double  *tkmg;          /* [SOI]    in   mineral thermal conductivity           */
    const double  *tkdry;         /* [SOI]    in   dry thermal conductivity               */
    const double  *csol;          /* [SOI]    in   solid heat capacity (J/m3/K)           */
    double        *eff_porosity;  /* [SOI]    inout (SoilWater reads; fractions write)    */
    double        *smp_l;         /* [SOI]    out  matric potential (mm)                  */
    double        *hk_l;          /* [SOI]    out  hydraulic conductivity (mm/s)          */
    double        *thk;           /* [SNO]    out  layer thermal conductivity             */
    /* SoilHydrologyType (soilhydrology_inst%) */
    double        *icefrac;       /* [SOI]    inout                                       */
    double        *qcharge;       /* (c)      out  aquifer recharge (mm/s)                */
    double        *num_substeps;  /* (c)      out  adaptive sub-step count (real, as ref) */
    double        *qin;           /* [SOI]    out  flux into layer (mm/s)                 */
    double        *qout;          /* [SOI]    out  flux out of layer (mm/s)               */
    /* WaterStateBulkType */
    double        *h2osoi_liq;    /* [SNO]    inout                                       */
    double        *h2osoi_ice;    /* [SNO]    inout                                       */
    double        *h2osfc;        /* (c)      inout surface water (mm)                    */
    double        *h2osno_no_layers; /* (c)   inout                                       */
    double        *int_snow;      /* (c)      inout                                       */
    /* WaterDiagnosticBulkType */
    const double  *frac_sno_eff;  /* (c)      in                                          */
    const double  *frac_h2osfc;   /* (c)      in                                          */
    double        *snow_depth;    /* (c)      inout                                       */
    double        *bw;            /* [SNOW_ONLY] out                                      */
    double        *snomelt_accum; /* (c)      inout                                       */
    /* WaterFluxBulkType */
    const double  *qflx_infl;     /* (c)      in   infiltration (mm/s)                    */
    const double  *qflx_rootsoi;  /* [SOI]    in   root sink (mm/s)                       */
    double        *qflx_snomelt;  /* (c)      out                                         */
    double        *qflx_snofrz;   /* (c)      out                                         */
    double        *qflx_snow_drain; /* (c)    out                                         */
    double        *qflx_snomelt_lyr; /* [SNOW_ONLY] out                                   */
    double        *qflx_snofrz_lyr;  /* [SNOW_ONLY] out                                   */
    double        *qflx_h2osfc_to_ice; /* (c) out                                         */
    /* TemperatureType */
    double        *t_soisno;      /* [SNO]    inout (K)                                   */
    double        *t_h2osfc;      /* (c)      inout                                       */
    double        *t_grnd;        /* (c)      inout                                       */
    double        *fact;          /* [SNO]    out                                         */
    double        *xmf;           /* (c)      out                                         */
    double        *xmf_h2osfc;    /* (c)      out                                         */
    double        *c_h2osfc;      /* (c)      out                                         */
    int32_t       *imelt;         /* [SNO]    out                                         */
    const double  *emg;           /* (c)      in   ground emissivity                      */
    /* EnergyFluxType */
    const double  *htvp;          /* (c)      in                                          */
    const double  *eflx_bot;      /* (c)      in                                          */
    double        *eflx_fgr12;    /* (c)      out                                         */
    double        *eflx_fgr;      /* [SOI]    out                                         */
    double        *eflx_snomelt;  /* (c)      out                                         */
    double        *eflx_h2osfc_to_snow; /* (c) out                                        */
    /* atm2lnd (downscaled to column) */
    const double  *forc_lwrad;    /* (c)      in                                          */
    /* SoilHydrologyType / SoilStateType, read by SoilWater only with the
     * water-table bottom boundary (lower_boundary_condition = CTSM_BC_WATERTABLE) */
    const double  *zwt;           /* (c)      in   water table depth (m)                  */
    const double  *smpmin;        /* (c)      in   minimum soil potential (mm)            */
} ctsm_col_state;

/* ---- patch state (PatchType / SolarAbsorbedType / EnergyFluxType) -------- */
typedef struct ctsm_patch_state {
    const int32_t *column;        /* (p) in  patch%column                                 */
    const double  *wtcol;         /* (p) in  patch%wtcol                                  */
    const int32_t *frac_veg_nosno;/* (p) in                                              */
    const double  *sabg;          /* (p) in                                              */
    const double  *sabg_soil;     /* (p) in                                              */
    const double  *sabg_snow;     /* (p) in                                              */
    const double  *sabg_lyr;      /* [LYR] in                                            */
    const double  *dlrad;         /* (p) in                                              */
    const double  *cgrnd;         /* (p) in                                              */
    const double  *eflx_sh_grnd;  /* (p) in                                              */
    const double  *eflx_sh_snow;  /* (p) in                                              */
    const double  *eflx_sh_soil;  /* (p) in                                              */
    const double  *eflx_sh_h2osfc;/* (p) in                                              */
    const double  *qflx_evap_soi; /* (p) in                                              */
    const double  *qflx_ev_snow;  /* (p) in                                              */
    const double  *qflx_ev_soil;  /* (p) in                                              */
    const double  *qflx_ev_h2osfc;/* (p) in                                              */
    double        *eflx_gnet;     /* (p) out                                             */
    double        *dgnetdT;       /* (p) out                                             */
    double        *sabg_chk;      /* (p) out                                             */
} ctsm_patch_state;

/* ---- CanopyFluxes parameters ----------------------------------------------
 * Param-file scalars read once by CanopyFluxesMod readParams
 * (CanopyFluxesMod.F90:158-189), PhotosynthesisMod readParams
 * (PhotosynthesisMod.F90:479-560) and FrictionVelocityMod
 * (zetamaxstable, FrictionVelocityMod.F90:511-538). */
typedef struct ctsm_canopy_params {
    double cv;             /* canopy turbulent transfer coefficient (m/s^0.5)   */
    double a_coef;         /* bare-soil transfer coefficient a (Sakaguchi-Zeng)  */
    double a_exp;          /* bare-soil transfer coefficient exponent           */
    double csoilc;         /* dense-canopy soil drag coefficient                */
    double z_dl;           /* litter thickness (m)                              */
    double lai_dl;         /* plant litter area index                           */
    double wind_min;       /* minimum wind speed (m/s)                          */
    double zetamaxstable;  /* max zeta under stable conditions                  */
    double kc25_coef, ko25_coef, cp25_yr2000;    /* mol/mol                      */
    double kcha, koha, cpha;                     /* activation energies (J/mol)  */
    double vcmaxha, jmaxha, tpuha, lmrha;        /* activation energies (J/mol)  */
    double vcmaxhd, jmaxhd, tpuhd, lmrhd;        /* deactivation energies        */
    double lmrse;                                /* entropy term (J/mol/K)       */
    double fnr, act25;                           /* Rubisco N ratio and activity */
    double theta_psii, theta_ip, fnps;           /* curvatures, fraction of PAR  */
    double tpu25ratio, kp25ratio;
    double vcmaxse_sf, jmaxse_sf, tpuse_sf, jmax25top_sf;
    double leaf_mr_vcm;    /* canopystate_inst%leaf_mr_vcm                       */
} ctsm_canopy_params;

/* Per-PFT parameters (pftconMod.F90), host arrays of npft rows indexed by
 * patch%itype (0-based PFT number, as in the reference).  The device table
 * holds CTSM_MAX_PFT rows: the reference's PFTs take the first ones and an
 * ensemble server gives each resident perturbed member private rows after
 * them (ctsm_set_pft_rows, itype renumbered like the column indices). */
#define CTSM_MAX_PFT 2048
typedef struct ctsm_pft_params {
    int32_t npft;
    int32_t pad;
    const double *dleaf, *c3psn, *slatop, *leafcn, *flnr, *fnitr;
    const double *mbbopt, *medlynslope, *medlynintercept;
    const double *smpso, *smpsc, *theta_cj;
    const double *z0v_Cr, *z0v_Cs, *z0v_c, *z0v_cw, *z0v_LAImax;
    const int32_t *is_tree, *is_shrub;
    const double *dbh, *fbw, *nstem, *rstem_per_dbh, *wood_density;
} ctsm_pft_params;

/* ---- canopy state: patch (p), column (c) and gridcell (g) arrays read or
 * written by CanopyFluxes (CanopyFluxesMod.F90:450-640 associate list), its
 * callees and the root-sink routine.  [SOIP] = (begp:endp, 1:nlevgrnd),
 * patch index fastest; [SOI] as in ctsm_col_state.  Patch outputs that
 * SoilTemperature reads (dlrad, cgrnd, eflx_sh_*, qflx_ev_*, qflx_evap_soi)
 * are the same arrays ctsm_patch_state names. */
typedef struct ctsm_canopy_state {
    /* PatchType */
    const int32_t *itype;          /* (p) in  PFT number                               */
    const int32_t *column;         /* (p) in  patch%column                             */
    const int32_t *gridcell;       /* (p) in  patch%gridcell                           */
    const double  *wtcol;          /* (p) in  patch%wtcol                              */
    const int32_t *frac_veg_nosno; /* (p) in                                           */
    const int32_t *active;         /* (p) in  patch%active (root sink, p2c)            */
    /* atm2lnd / GridcellType (gridcell) */
    const double  *forc_u;         /* (g) in                                           */
    const double  *forc_v;         /* (g) in                                           */
    const double  *forc_pco2;      /* (g) in  Pa                                       */
    const double  *forc_po2;       /* (g) in  Pa                                       */
    const double  *forc_hgt_u;     /* (g) in  m                                        */
    const double  *forc_hgt_t;     /* (g) in                                           */
    const double  *forc_hgt_q;     /* (g) in                                           */
    const double  *dayl;           /* (g) in  s                                        */
    const double  *max_dayl;       /* (g) in  s                                        */
    const int32_t *near_local_noon;/* (g) in  is_near_local_noon(londeg, 3600)          */
    /* column forcing and ground state */
    const double  *forc_q;         /* (c) in                                           */
    const double  *forc_pbot;      /* (c) in                                           */
    const double  *forc_th;        /* (c) in                                           */
    const double  *forc_rho;       /* (c) in                                           */
    const double  *thv;            /* (c) in                                           */
    const double  *qg;             /* (c) in                                           */
    const double  *qg_snow;        /* (c) in                                           */
    const double  *qg_soil;        /* (c) in                                           */
    const double  *qg_h2osfc;      /* (c) in                                           */
    const double  *dqgdT;          /* (c) in                                           */
    const double  *z0mg;           /* (c) in                                           */
    const double  *soilbeta;       /* (c) in                                           */
    const double  *soilresis;      /* (c) in                                           */
    double        *h2osoi_liqvol;  /* [SOI] out                                        */
    double        *qflx_tran_veg_col; /* (c) out (p2c, clm_driver.F90:1715)             */
    double        *rootr_col;      /* [SOI] out root sink weights                      */
    double        *qflx_rootsoi_col; /* [SOI] out (same array ctsm_col_state reads)   */
    /* canopy state inputs (p) */
    const double  *elai;
    const double  *esai;
    const double  *laisun;
    const double  *laisha;
    const double  *htop;
    const double  *sabv;
    const double  *parsun_z;       /* (p) nlevcan = 1                                  */
    const double  *parsha_z;
    const double  *laisun_z;
    const double  *laisha_z;
    const double  *vcmaxcintsun;
    const double  *vcmaxcintsha;
    const int32_t *nrad;
    const double  *thm;
    const double  *emv;
    const double  *fwet;
    const double  *fdry;
    const double  *t_a10;
    const double  *o3coefvsun;
    const double  *o3coefvsha;
    const double  *o3coefgsun;
    const double  *o3coefgsha;
    const double  *rootfr;         /* [SOIP] in                                        */
    /* inout (p) */
    double        *displa;
    double        *z0mv;
    double        *t_veg;
    double        *t_stem;
    double        *snocan;
    double        *liqcan;
    double        *cgrnds;
    double        *cgrndl;
    /* outputs (p): frictionvel_inst */
    double        *z0hv, *z0qv;
    double        *forc_hgt_u_patch, *forc_hgt_t_patch, *forc_hgt_q_patch;
    double        *ram1, *rb1, *rah1, *rah2, *raw1, *raw2;
    double        *ustar, *um, *uaf, *taf, *qaf, *obu, *zeta, *vpd, *num_iter;
    double        *u10, *u10_clm, *va, *fv, *vds;
    /* energyflux_inst */
    double        *btran;
    double        *rresis;         /* [SOIP] out                                       */
    double        *rootr;          /* [SOIP] out                                       */
    double        *eflx_sh_veg, *eflx_sh_stem, *eflx_sh_grnd, *eflx_sh_snow, *eflx_sh_soil, *eflx_sh_h2osfc;
    double        *taux, *tauy, *dlrad, *ulrad, *cgrnd, *dhsdt_canopy;
    /* waterflux / waterdiagnostic */
    double        *qflx_evap_soi, *qflx_ev_snow, *qflx_ev_soil, *qflx_ev_h2osfc;
    double        *qflx_tran_veg, *qflx_evap_veg;
    double        *rh_af, *q_ref2m, *rh_ref2m, *rh_ref2m_r, *vpd_ref2m, *iwue_ln;
    /* temperature / canopystate */
    double        *t_ref2m, *t_ref2m_r, *t_skin;
    double        *stem_biomass, *leaf_biomass, *dleaf_patch;
    /* photosyns_inst (nlevcan = 1: the _z arrays are (p)) */
    double        *rssun, *rssha, *psnsun, *psnsha;
    double        *psnsun_wc, *psnsun_wj, *psnsun_wp, *psnsha_wc, *psnsha_wj, *psnsha_wp;
    double        *fpsn, *fpsn_wc, *fpsn_wj, *fpsn_wp;
    double        *lmrsun, *lmrsha, *gs_mol_sun, *gs_mol_sha, *gs_mol_sun_ln, *gs_mol_sha_ln;
    double        *cisun_z, *cisha_z, *rssun_z, *rssha_z, *lmrsun_z, *lmrsha_z, *psnsun_z, *psnsha_z;
    double        *vcmax_z, *tpu_z, *kp_z, *ac, *aj, *ap, *ag, *an, *gb_mol, *gs_mol;
    double        *cp, *kc, *ko, *qe, *bbb, *mbb, *rh_leaf, *vpd_can, *lnca;
    int32_t       *c3flag;
    /* diagnostic: the canopy energy balance residual err(p) that the
     * reference checks against 0.1 W/m2 (CanopyFluxesMod.F90:1742-1756) */
    double        *canopy_err;
} ctsm_canopy_state;

/* ---- SoilFluxes state (SoilFluxesMod.F90:38-518) --------------------------
 * The patch and column arrays SoilFluxes reads and writes beyond
 * ctsm_col_state (column fields: snl, patchi/patchf, frac_sno_eff,
 * frac_h2osfc, h2osoi_liq/ice, t_soisno, t_h2osfc, t_grnd, emg, forc_lwrad,
 * htvp, xmf, xmf_h2osfc, fact, c_h2osfc, eflx_h2osfc_to_snow) and
 * ctsm_patch_state (column, wtcol, frac_veg_nosno, sabg_soil, sabg_snow,
 * dlrad).  Patch arrays are (begp:endp), column arrays (begc:endc), t_ssbef
 * is a [SNO] array like t_soisno.  Non-urban soil / crop landunits. */
typedef struct ctsm_soilflux_state {
    const int32_t *active;        /* (p) in  patch%active (the p2c of errsoi)            */
    /* TemperatureType: saved at the start of the step (BiogeophysPreFluxCalcsMod.F90:302-331) */
    double        *t_ssbef;       /* [SNO] in (written by the land step's save)          */
    double        *t_h2osfc_bef;  /* (c)   in (written by the land step's save)          */
    /* corrected to the new ground temperature (inout) */
    double        *eflx_sh_grnd, *qflx_evap_soi, *qflx_ev_snow, *qflx_ev_soil, *qflx_ev_h2osfc;
    /* CanopyFluxes outputs read here */
    const double  *cgrnds, *cgrndl, *ulrad, *eflx_sh_veg, *eflx_sh_stem, *qflx_evap_veg, *qflx_tran_veg;
    /* outputs */
    double        *eflx_soil_grnd, *eflx_soil_grnd_r, *eflx_sh_tot, *eflx_sh_tot_r;
    double        *qflx_evap_tot, *eflx_lh_tot, *eflx_lh_tot_r;
    double        *qflx_evap_can, *eflx_lh_vege, *eflx_lh_vegt, *eflx_lh_grnd;
    double        *eflx_lwrad_out, *eflx_lwrad_out_r, *eflx_lwrad_net, *eflx_lwrad_net_r, *t_skin;
    double        *qflx_liqevap_from_top_layer, *qflx_solidevap_from_top_layer;
    double        *qflx_liqdew_to_top_layer, *qflx_soliddew_to_top_layer;
    double        *errsoi_patch;  /* (p) soil energy conservation error (W/m2)           */
    double        *errsoi_col;    /* (c) p2c of errsoi_patch (W/m2)                      */
} ctsm_soilflux_state;

/* ---- surface water inputs of the soil (HydrologyNoDrainageMod.F90:306-336) --
 * Column arrays (begc:endc) of SaturatedExcessRunoff (TOPMODEL fsat),
 * SetQflxInputs, InfiltrationExcessRunoff (qinmax from hksat), RouteInfiltrationExcess,
 * UpdateH2osfc, Infiltration and TotalSurfaceRunoff for soil / crop columns;
 * snl, frac_sno_eff, frac_h2osfc, icefrac and hksat come from ctsm_col_state.
 * qflx_infl is the SoilWater input (ctsm_col_state.qflx_infl) and h2osfc the
 * surface water (ctsm_col_state.h2osfc): pass the same arrays. */
typedef struct ctsm_surface_water_state {
    /* inputs */
    const double  *qflx_rain_plus_snomelt, *qflx_snow_h2osfc, *qflx_floodc;
    /* column values of the patch fluxes (p2c); ctsm_land_step_full_dev forms
     * them from the patches, as clm_drv_patch2col does */
    double        *qflx_liqevap_from_top_layer, *qflx_ev_soil, *qflx_ev_h2osfc;
    const double  *zwt, *zwt_perched, *frost_table, *wtfact;                        /* SoilHydrologyType, SoilStateType */
    const double  *topo_slope, *h2osfc_thresh, *frac_h2osfc_nosnow;
    /* inout */
    double        *h2osfc;
    /* outputs */
    double        *fsat, *fcov, *qflx_sat_excess_surf, *qflx_top_soil, *qflx_in_soil, *qflx_top_soil_to_h2osfc;
    double        *qinmax, *qflx_infl_excess, *qflx_in_soil_limited, *qflx_in_h2osfc, *qflx_infl_excess_surf;
    double        *qflx_h2osfc_surf, *qflx_h2osfc_drain, *qflx_infl, *qflx_surf;
} ctsm_surface_water_state;

/* ---- water tables and condensation after SoilWater (HydrologyNoDrainageMod.F90:
 * 355-373, the CLM5/6 path without the aquifer layer) -------------------------
 * Column arrays of PerchedWaterTable, ThetaBasedWaterTable and RenewCondensation
 * for soil / crop columns; t_soisno, dz, z, zi, watsat, nbedrock, snl,
 * frac_h2osfc, h2osoi_liq and h2osoi_ice come from ctsm_col_state.  zwt,
 * zwt_perched and frost_table are the surface water chain's inputs of the next
 * step: pass the same arrays as in ctsm_surface_water_state. */
typedef struct ctsm_water_table_state {
    double        *zwt;                          /* (c)   inout water table depth (m)           */
    double        *zwt_perched, *frost_table;    /* (c)   out   (m)                             */
    double        *h2osoi_vol;                   /* [SOI] out   for the layers the two searches
                                                    visit, as the reference; may be NULL        */
    /* column values of the patch fluxes (p2c, mm H2O/s); ctsm_land_step_full_dev
     * forms them from the patches, as clm_drv_patch2col does */
    double        *qflx_liqdew_to_top_layer, *qflx_soliddew_to_top_layer, *qflx_solidevap_from_top_layer;
} ctsm_water_table_state;

/* ---- drainage (HydrologyDrainageMod.F90:127-245 without the aquifer layer) --
 * Column arrays of PerchedLateralFlow (SoilHydrologyMod.F90:1645-1929) and
 * SubsurfaceLateralFlow (:2029-2565) for soil / crop columns off hillslopes, and
 * the runoff sum of HydrologyDrainage; dz, z, zi, watsat, bsw, sucsat, hksat,
 * eff_porosity, nbedrock, h2osoi_liq, h2osoi_ice, h2osfc and icefrac (rewritten
 * from the ice, as the reference does) come from ctsm_col_state.  zwt,
 * zwt_perched, frost_table, topo_slope and qflx_surf are the same arrays as in
 * the surface water and water-table states. */
typedef struct ctsm_drainage_state {
    const double  *topo_slope, *zwt_perched, *frost_table;   /* (c) in  */
    const double  *qflx_surf, *qflx_snwcp_liq;               /* (c) in  (mm H2O/s) */
    double        *zwt;                                      /* (c) inout (m) */
    double        *h2osoi_vol;                               /* [SOI] out, may be NULL */
    /* outputs (mm H2O/s) */
    double        *qflx_drain_perched, *qflx_latflow_out, *qflx_latflow_in, *qflx_rsub_sat;
    double        *qflx_ice_runoff_xs, *qflx_drain, *qflx_qrgwl, *qflx_runoff;
} ctsm_drainage_state;

/* Param-file / namelist scalars of the hydrology: fff (SaturatedExcessRunoffMod
 * readParams), pc and mu (SurfaceWaterMod readParams), aq_sp_yield_min,
 * n_baseflow, baseflow_scalar and perched_baseflow_scalar (SoilHydrologyMod
 * readParams), h2osfcflag (SoilHydrologyType, 1 = surface water active). */
typedef struct ctsm_hydro_params {
    double  fff, pc, mu;
    double  aq_sp_yield_min, n_baseflow, baseflow_scalar, perched_baseflow_scalar;
    int32_t h2osfcflag;
    int32_t pad;
} ctsm_hydro_params;

/* ---- plant hydraulic stress, soil side (use_hydrstress = .true.) -----------
 * Arrays of the root-soil interface conductance block of
 * PhotosynthesisHydraulicStress (PhotosynthesisMod.F90:3062-3114) and of
 * Compute_EffecRootFrac_And_VertTranSink_HydStress (SoilWaterPlantSinkMod.F90:
 * 236-328).  Members that CTSM keeps in the column and patch types (itype,
 * column, active, frac_veg_nosno, wtcol, rootfr, patchi/patchf, dz, z, hksat,
 * hk_l, smp_l, qflx_rootsoi) are the same arrays as in ctsm_col_state /
 * ctsm_canopy_state.  [VEG] arrays are (begp:endp, 1:4) with the patch
 * fastest: vegwp(p, k) for the segments sun, sha, xyl, root (k = 1..4). */
typedef struct ctsm_phs_state {
    /* patch */
    const int32_t *itype, *column, *active, *frac_veg_nosno;   /* (p) */
    const double  *wtcol;                                      /* (p) */
    const double  *froot_carbon;                               /* (p) fine root carbon (gC/m2) */
    const double  *tsai, *tlai;                                /* (p) stem / leaf area index, no snow burial */
    const double  *rootfr;                                     /* [SOIP] root fraction */
    const double  *vegwp;                                      /* [VEG] vegetation water potential (mm) */
    double        *k_soil_root;                                /* [SOIP] out soil-root conductance (1/s) */
    double        *root_conductance, *soil_conductance;        /* [SOIP] out */
    double        *qflx_hydr_redist;                           /* (p) out negative root-soil flux (mm H2O/s) */
    /* column */
    const int32_t *patchi, *patchf;                            /* (c) */
    const double  *dz, *z;                                     /* [SNO] (m) */
    const double  *hksat, *hk_l, *smp_l;                       /* [SOI] (mm/s, mm/s, mm) */
    double        *qflx_rootsoi;                               /* [SOI] out root-soil exchange (mm H2O/s) */
    double        *qflx_phs_neg;                               /* (c) out sum of negative layer exchanges */
} ctsm_phs_state;

/* PHS parameters of the param file (PhotosynthesisMod readParams) and pftcon,
 * host arrays indexed by the vegetation type (0..npft-1): krmax (root
 * conductance per area per length), root_radius (m), root_density (g biomass /
 * m3 root), froot_leaf (fine root to leaf ratio); psi50 (mm) and ck of the
 * Weibull vulnerability curve, segment-major: psi50[(k-1)*npft + ivt]. */
typedef struct ctsm_phs_params {
    int32_t       npft;
    int32_t       pad;
    const double *krmax, *root_radius, *root_density, *froot_leaf;
    const double *psi50, *ck;
    const double *kmax;   /* params_inst%kmax, segment-major as psi50; may be NULL when
                             neither ctsm_phs_calcstress nor the PHS canopy path is used */
    const double *crop;   /* pftcon%crop (0 or 1; PhotosynthesisMod.F90:3516, 3739); may be
                             NULL: no crop types */
} ctsm_phs_params;

/* Per-patch arrays of calcstress (PhotosynthesisMod.F90:4490-4710), the
 * vegetation water potential solve of plant hydraulic stress: the four-segment
 * Newton iteration (spacF / spacA, :4715-4976) from the given vegwp, the
 * stomatal conductances it allows (getqflx, :5080-5164), getvegwp (:4979-5077)
 * where Newton stops on a singular Jacobian or there is no transpiration.
 * gs_mol_sun / gs_mol_sha are the unstressed conductances the caller's ci
 * solve produced; bsun / bsha the resulting stress factors.  A positive
 * vegwp(sun) on input marks night (the reference's convention): the night
 * branch then also writes qflx_tran_veg, and vegwp_pd where local_morning is
 * set (the caller's get_local_time(londeg) < isecspday/2). */
typedef struct ctsm_phs_stress_state {
    /* patch */
    const int32_t *itype, *column, *local_morning;              /* (p) */
    const double  *laisun, *laisha, *elai, *esai, *tsai, *htop;  /* (p) */
    const double  *fdry, *thm;                                  /* (p) */
    const double  *gb_mol, *gs_mol_sun, *gs_mol_sha;            /* (p) umol H2O/m2/s */
    const double  *qsatl, *qaf;                                 /* (p) kg/kg */
    const double  *k_soil_root;                                 /* [SOIP] */
    double        *vegwp;                                       /* [VEG] inout (mm) */
    double        *vegwp_pd;                                    /* [VEG] out (mm, spval unless night morning) */
    double        *bsun, *bsha;                                 /* (p) out */
    double        *qflx_tran_veg;                               /* (p) out at night (mm H2O/s) */
    /* column */
    const double  *forc_rho, *forc_pbot;                        /* (c) */
    const double  *z;                                           /* [SNO] (m) */
    const double  *smp_l;                                       /* [SOI] (mm) */
} ctsm_phs_stress_state;

/* CanopyFluxes with plant hydraulic stress (use_hydrstress = 1):
 * PhotosynthesisHydraulicStress (PhotosynthesisMod.F90:2704-3811) with the
 * coupled sun / shade ci solve hybrid_PHS, brent_PHS, ci_func_PHS
 * (:3815-4486) and calcstress inside it, in every iteration of the leaf
 * temperature, and the PHS branches of CanopyFluxes (CanopyFluxesMod.F90:
 * 1133-1166, 1219-1230, 1353-1356).  The arrays of that path beyond
 * ctsm_canopy_state and ctsm_phs_state: [VEG] arrays (begp:endp, 1:4) with
 * the patch fastest (sun, sha, xyl, root); [PHS2] arrays (begp:endp, 1:2)
 * with the patch fastest (sun, sha) for the single canopy layer (nlevcan =
 * 1).  Photosynthesis outputs both paths share (psnsun ..., rssun, rssha,
 * cisun_z, gs_mol_sun / _sha and their noon values, gb_mol, kc, ko, cp, qe,
 * bbb, mbb, lnca, vpd_can, btran, qflx_tran_veg) stay in ctsm_canopy_state;
 * the PHS path leaves the non-PHS arrays vcmax_z, tpu_z, kp_z, ac, aj, ap,
 * ag, an, gs_mol and rh_leaf untouched, as the reference does. */
typedef struct ctsm_phs_canopy_state {
    /* patch */
    const double  *tlai_z;             /* (p) tlai_z(p, 1), the layer's leaf area */
    const int32_t *local_morning;      /* (p) nonzero where a night calcstress writes vegwp_pd */
    double        *vegwp;              /* [VEG] inout (mm) */
    double        *vegwp_ln, *vegwp_pd;  /* [VEG] out (mm; spval away from noon / morning) */
    double        *bsun, *bsha;        /* (p) inout (no canopy layer: btran = the old bsun) */
    double        *an_sun, *an_sha;    /* (p) out net photosynthesis (umol CO2/m2/s) */
    double        *vcmax_z_phs, *tpu_z_phs, *kp_z_phs;   /* [PHS2] out */
    double        *ac_phs, *aj_phs, *ap_phs, *ag_phs;    /* [PHS2] out */
} ctsm_phs_canopy_state;

/* ---- launch shapes (no counterpart in the reference: set by the autotuner) --
 * richards_lanes: lanes cooperating on one column in the Richards solve
 *   (1, 2, 4 or 8), or 0: chosen per wavefront from the deepest bedrock index
 *   among 32 consecutive columns (3 layers per lane: 2 lanes up to 6 layers,
 *   4 up to 12, 8 up to 20); the heat solve always runs one column per lane.
 * fused_single_kernel: 1 = the whole soil step in one kernel (one lane per
 *   column), 0 = heat solve kernel + Richards kernel (2 launches per step).
 * canopy_first_iterations, canopy_round_iterations: CanopyFluxes runs its
 *   leaf-temperature iteration in rounds; the first round takes every patch of
 *   the filter for at most canopy_first_iterations iterations, each later round
 *   only the patches still iterating (compacted) for at most
 *   canopy_round_iterations more.  canopy_first_iterations > itmax_canopy_fluxes
 *   runs the whole iteration in one pass.  Both >= 1; results do not depend on
 *   them.
 * step_streams: the fused soil step splits its filter into this many
 *   contiguous parts run on as many streams of the context at once, joined
 *   before the call's later work (1..4).  A batch stored in work-binned
 *   order keeps the parts balanced by dealing the bins over the parts.
 * plain_routines: 1 = the land steps (ctsm_land_step_full_dev and its
 *   captures) run the plain form: SoilTemperature and SoilWater as one launch
 *   per reference routine with intermediates through HBM (the plain kernels of
 *   the *_dev entries' `plain` argument), CanopyFluxes as one launch that
 *   iterates every patch to convergence (no compacted rounds).  The comparison
 *   baseline of the batched step; results are the same.
 * canopy_waves: the register budget of the CanopyFluxes kernel, in waves per
 *   SIMD (1, 2 or 3: one compiled instance each; a larger budget spills more
 *   to scratch).  Results do not depend on it. */
typedef struct ctsm_tuning {
    int32_t richards_lanes;
    int32_t richards_block;
    int32_t soiltemp_block;
    int32_t fused_single_kernel;
    int32_t canopy_first_iterations;
    int32_t canopy_round_iterations;
    int32_t step_streams;
    int32_t plain_routines;
    int32_t canopy_waves;
} ctsm_tuning;

typedef struct ctsm_ctx ctsm_ctx;

/* ---- set-up --------------------------------------------------------------- */

/* Reference defaults of the CLM6 "standard" structure (namelist_defaults_ctsm.xml:
 * 462-488, 511, 555-556; e_ice from the CLM5 tech note, Omega = 6). */
void ctsm_config_default(ctsm_config *cfg);

/* Creates a context on HIP device `device` (replaces init_soilwater_movement and
 * CanopyFluxesReadNML: the namelist block is copied and checked once).  Returns
 * CTSM_E_CONFIG for combinations the reference itself rejects or that this
 * library does not run (soilwater_movement_method != 1, lower BC other than
 * flux / zero flux, upper BC other than flux). */
int ctsm_create(const ctsm_config *cfg, int device, ctsm_ctx **out);
int ctsm_destroy(ctsm_ctx *ctx);

int ctsm_set_tuning(ctsm_ctx *ctx, const ctsm_tuning *t);
int ctsm_get_tuning(const ctsm_ctx *ctx, ctsm_tuning *t);

/* First failing column (reference 1-based index) and code of the last call. */
int ctsm_last_error(const ctsm_ctx *ctx, int32_t *code, int32_t *index);

/* Canopy parameters, copied to the device once (replaces the readParams of
 * CanopyFluxesMod / PhotosynthesisMod and the pftcon arrays).  Must precede any
 * canopy call; CTSM_E_ARG for npft outside 1..CTSM_MAX_PFT. */
int ctsm_set_canopy_params(ctsm_ctx *ctx, const ctsm_canopy_params *prm,
                           const ctsm_pft_params *pft);

/* ---- drop-in entry points (host arrays, synchronous) ----------------------
 * Each copies the filtered columns' fields to the device, runs the kernel and
 * copies the outputs back before returning, so the Fortran caller reads its
 * own arrays right after the call as it does today. */

/* Replaces SoilWater (SoilWaterMovementMod.F90:240-243) with
 * soilwater_movement_method = 1 (soilwater_moisture_form :976-1440). */
int ctsm_soilwater(ctsm_ctx *ctx, const ctsm_layout *lay,
                   const ctsm_col_state *col,
                   int32_t num_hydrologyc, const int32_t *filter_hydrologyc,
                   double dtime);

/* Replaces SoilTemperature (SoilTemperatureMod.F90:92-508) for non-urban,
 * non-lake columns without excess ice (use_excess_ice = .false.). */
int ctsm_soiltemperature(ctsm_ctx *ctx, const ctsm_layout *lay,
                         const ctsm_col_state *col, const ctsm_patch_state *pch,
                         int32_t num_nolakec, const int32_t *filter_nolakec,
                         double dtime);

/* Replaces SetSoilWaterFractions (SoilHydrologyMod.F90:202-256) with
 * excess_ice = 0: icefrac and eff_porosity from h2osoi_ice. */
int ctsm_soilwater_fractions(ctsm_ctx *ctx, const ctsm_layout *lay,
                             const ctsm_col_state *col,
                             int32_t num_hydrologyc, const int32_t *filter_hydrologyc);

/* Replaces CanopyFluxes (CanopyFluxesMod.F90:191-1765) without FATES, CN,
 * plant hydraulic stress, LUNA, C13/C14, methane conductances and human
 * stress indices, for nlevcan = 1; Photosynthesis / hybrid / brent / ci_func
 * (PhotosynthesisMod.F90:1243-2701), FrictionVelocity / MoninObukIni
 * (FrictionVelocityMod.F90:754-1209), QSat (QSatMod.F90:61-127) and
 * calc_root_moist_stress (SoilMoistStressMod.F90:312-514) run inside.  Patch
 * arrays are mirrored over the whole allocation (begp_alloc .. +ldp-1),
 * column and gridcell arrays likewise. */
int ctsm_canopyfluxes(ctsm_ctx *ctx, const ctsm_layout *lay,
                      const ctsm_col_state *col, const ctsm_canopy_state *can,
                      int32_t num_exposedvegp, const int32_t *filter_exposedvegp,
                      double dtime);

/* Replaces the p2c of qflx_tran_veg (clm_driver.F90:1714-1716, subgridAveMod
 * p2c_1d_filter) followed by Compute_EffecRootFrac_And_VertTranSink_Default
 * (SoilWaterPlantSinkMod.F90:332-424): rootr_col, qflx_rootsoi_col. */
int ctsm_root_sink(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                   const ctsm_canopy_state *can,
                   int32_t num_filterc, const int32_t *filterc);

/* Replaces SoilFluxes (SoilFluxesMod.F90:38-518, called at clm_driver.F90:921)
 * for soil / crop patches: fluxes corrected to the new ground temperature,
 * evaporation partitioned and limited, ground heat flux, total and history
 * fluxes, outgoing longwave, and the soil energy error errsoi (patch, then
 * p2c to column).  sf->t_ssbef and sf->t_h2osfc_bef hold the values saved
 * before the step. */
int ctsm_soilfluxes(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                    const ctsm_patch_state *pch, const ctsm_soilflux_state *sf,
                    int32_t num_nolakec, const int32_t *filter_nolakec,
                    int32_t num_nolakep, const int32_t *filter_nolakep, double dtime);

/* The soil energy part of EnergyBalanceCheck (BalanceCheckMod.F90:1101-1119):
 * max |errsoi_col| over the filtered columns and the lowest column index
 * holding it (maxloc).  The caller applies the reference's thresholds (warn
 * above 1e-5 W/m2, endrun above 1e-4 W/m2 after the skip steps). */
int ctsm_soil_balance_check(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_soilflux_state *sf,
                            int32_t num_filterc, const int32_t *filterc,
                            double *errsoi_max, int32_t *index);

/* Replaces the surface water input chain of HydrologyNoDrainage
 * (HydrologyNoDrainageMod.F90:306-336: SaturatedExcessRunoff with the TOPMODEL
 * fsat, SetQflxInputs, InfiltrationExcessRunoff with qinmax from hksat,
 * RouteInfiltrationExcess, UpdateH2osfc, Infiltration, TotalSurfaceRunoff) for
 * soil / crop columns; produces qflx_infl for SoilWater.  SetFloodc is the
 * caller's (qflx_floodc is an input). */
int ctsm_surface_water(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_surface_water_state *sw, const ctsm_hydro_params *prm,
                       int32_t num_hydrologyc, const int32_t *filter_hydrologyc, double dtime);

/* Replaces the calls after SoilWater in HydrologyNoDrainage
 * (HydrologyNoDrainageMod.F90:355-373 without the aquifer layer):
 * PerchedWaterTable (frost table, perched water table), ThetaBasedWaterTable
 * (water table) and RenewCondensation (dew, frost and sublimation of the top
 * soil layer of snow-free columns) for soil / crop columns.  Returns
 * CTSM_E_ENDRUN with the column, as the reference's endrun, when sublimation
 * leaves the top layer's ice significantly negative. */
int ctsm_water_table(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                     const ctsm_water_table_state *wt, int32_t num_hydrologyc,
                     const int32_t *filter_hydrologyc, double dtime);

/* Replaces the soil-column part of HydrologyDrainage without the aquifer layer
 * or hillslopes (HydrologyDrainageMod.F90:137-149, 160-245): PerchedLateralFlow,
 * SubsurfaceLateralFlow (baseflow from the water table, the water table's rise
 * or fall, saturation excess pushed up, the minimum-water corrections), the
 * volumetric water and qflx_runoff. */
int ctsm_drainage(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                  const ctsm_drainage_state *dr, const ctsm_hydro_params *prm,
                  int32_t num_hydrologyc, const int32_t *filter_hydrologyc, double dtime);

/* Stores the PHS parameters in the context (copied to the device; replaces
 * params_inst%krmax / psi50 / ck and pftcon%root_radius / root_density /
 * froot_leaf, PhotosynthesisMod.F90:3062-3114 and :5167-5195).  Must precede
 * the two calls below; CTSM_E_ARG for npft outside 1..CTSM_MAX_PFT or a NULL
 * table. */
int ctsm_set_phs_params(ctsm_ctx *ctx, const ctsm_phs_params *prm);

/* Replaces the root-soil interface conductance loop at the head of
 * PhotosynthesisHydraulicStress (PhotosynthesisMod.F90:3062-3114): per patch of
 * the filter and soil layer, the Gardner soil conductance, the root conductance
 * attenuated by the root segment's vulnerability curve plc(smp) (:5167-5195),
 * and their series conductance k_soil_root (zero in the top layer and where
 * there are no roots). */
int ctsm_phs_soil_root_conductance(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs,
                                   int32_t num_filterp, const int32_t *filterp);

/* Replaces Compute_EffecRootFrac_And_VertTranSink_HydStress
 * (SoilWaterPlantSinkMod.F90:236-328): per column of the filter and soil layer
 * the patch-weighted root-soil exchange k_soil_root (smp - vegwp(root) - z),
 * its negative part per column and per patch. */
int ctsm_phs_root_sink(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs,
                       int32_t num_filterc, const int32_t *filterc);

/* calcstress (PhotosynthesisMod.F90:4490-4710) for every patch of the filter
 * at once: the reference calls it patch by patch from ci_func_PHS (:4227-4486);
 * a caller with the ci solve batched calls this once per evaluation round.
 * Needs ctsm_set_phs_params with kmax. */
int ctsm_phs_calcstress(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_stress_state *st,
                        int32_t num_filterp, const int32_t *filterp);

/* ---- device-resident entry points -----------------------------------------
 * Same arguments, but every pointer (fields and filter) is device memory and
 * the call only enqueues work on the context's stream.  filter == NULL means
 * every column of the layout, in order.  `plain` selects the unfused form
 * (one launch per reference routine, intermediates through HBM) used as the
 * comparison baseline; plain = 0 runs the fused per-column kernels. */
int ctsm_soilwater_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       int32_t num_filter, const int32_t *filter_dev, double dtime,
                       int32_t plain);
int ctsm_soiltemperature_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                             const ctsm_col_state *col, const ctsm_patch_state *pch,
                             int32_t num_filter, const int32_t *filter_dev, double dtime,
                             int32_t plain);
int ctsm_soilwater_fractions_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                                 const ctsm_col_state *col,
                                 int32_t num_filter, const int32_t *filter_dev);

/* One full soil step per column in ONE launch: SoilTemperature, then
 * SetSoilWaterFractions, then SoilWater -- the order clm_drv runs them in
 * (clm_driver.F90:900, HydrologyNoDrainageMod.F90:302-346). */
int ctsm_soil_step_fused_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                             const ctsm_col_state *col, const ctsm_patch_state *pch,
                             int32_t num_filter, const int32_t *filter_dev, double dtime);
/* nsteps such steps in one call.  With step_streams > 1 each part of the
 * columns runs all nsteps on its own stream between one fork and one join
 * (the parts need not wait for each other between steps). */
int ctsm_soil_steps_fused_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                              const ctsm_col_state *col, const ctsm_patch_state *pch,
                              int32_t num_filter, const int32_t *filter_dev, double dtime,
                              int32_t nsteps);

int ctsm_canopyfluxes_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                          const ctsm_col_state *col, const ctsm_canopy_state *can,
                          int32_t num_filterp, const int32_t *filterp_dev, double dtime);
/* ctsm_canopyfluxes_dev without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc after set-up).
 * The compacted rounds' iteration state, patch lists and counters live in a
 * fixed device array of CTSM_NOALLOC_SCRATCH_BYTES that the library holds on
 * every GPU it runs on until the process ends, instead of the context's
 * growable buffer; nothing is allocated or freed on the call path.  Results
 * equal ctsm_canopyfluxes_dev's.  A call that would need more scratch
 * (canopy_work_bytes: 136 B per patch of the filter plus 4 B per round) returns
 * CTSM_E_SCRATCH before launching anything, ctsm_last_error's index then being
 * the KiB it needs; 128 MiB holds about 986,000 patches.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two contexts,
 * two streams) race, so serialise them. */
#define CTSM_NOALLOC_SCRATCH_BYTES (128u << 20)
int ctsm_canopyfluxes_dev_noalloc(ctsm_ctx *ctx, const ctsm_layout *lay,
                                  const ctsm_col_state *col, const ctsm_canopy_state *can,
                                  int32_t num_filterp, const int32_t *filterp_dev, double dtime);
int ctsm_root_sink_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_canopy_state *can,
                       int32_t num_filterc, const int32_t *filterc_dev);

int ctsm_surface_water_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                           const ctsm_surface_water_state *sw, const ctsm_hydro_params *prm,
                           int32_t num_filterc, const int32_t *filterc_dev, double dtime);
int ctsm_water_table_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                         const ctsm_water_table_state *wt, int32_t num_filterc, const int32_t *filterc_dev,
                         double dtime);
int ctsm_drainage_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                      const ctsm_drainage_state *dr, const ctsm_hydro_params *prm,
                      int32_t num_filterc, const int32_t *filterc_dev, double dtime);
int ctsm_phs_soil_root_conductance_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs,
                                       int32_t num_filterp, const int32_t *filterp_dev);
int ctsm_phs_root_sink_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs,
                           int32_t num_filterc, const int32_t *filterc_dev);
/* CanopyFluxes with plant hydraulic stress (CanopyFluxesMod.F90:191-1765 with
 * use_hydrstress = .true.; the context's config has use_hydrstress = 1 and
 * ctsm_set_phs_params was given kmax): the root-soil conductances of the
 * filter's patches into phs (k_soil_root, root_conductance,
 * soil_conductance: the loop at the head of PhotosynthesisHydraulicStress,
 * :3062-3114, whose inputs do not change during the iteration), then the
 * leaf-temperature iteration with PhotosynthesisHydraulicStress in it.  phs
 * supplies the conductance members and z / smp_l; col, can as for
 * ctsm_canopyfluxes_dev.  Errors as ctsm_canopyfluxes_dev, and
 * CTSM_E_ENDRUN for a brent_PHS root not bracketed or a negative stomatal
 * conductance (the reference's endrun). */
int ctsm_canopyfluxes_phs_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                              const ctsm_canopy_state *can, const ctsm_phs_state *phs,
                              const ctsm_phs_canopy_state *phc, int32_t num_filterp, const int32_t *filterp_dev,
                              double dtime);
/* Drop-in form (host arrays, as ctsm_canopyfluxes): replaces the call of
 * CanopyFluxes at clm_driver.F90:766 when use_hydrstress = .true.  Copies the
 * members in, runs ctsm_canopyfluxes_phs_dev's work, copies the outputs
 * (col, can, the three conductances of phs, every phc member but tlai_z and
 * local_morning) back and waits. */
int ctsm_canopyfluxes_phs(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                          const ctsm_canopy_state *can, const ctsm_phs_state *phs,
                          const ctsm_phs_canopy_state *phc, int32_t num_exposedvegp, const int32_t *filter_exposedvegp,
                          double dtime);
int ctsm_phs_calcstress_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_stress_state *st,
                            int32_t num_filterp, const int32_t *filterp_dev);
int ctsm_soilfluxes_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                        const ctsm_patch_state *pch, const ctsm_soilflux_state *sf,
                        int32_t num_filterc, const int32_t *filterc_dev,
                        int32_t num_filterp, const int32_t *filterp_dev, double dtime);
/* Device form of ctsm_soil_balance_check: waits for the stream and returns the
 * maximum and its column on the host. */
int ctsm_soil_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_soilflux_state *sf,
                                int32_t num_filterc, const int32_t *filterc_dev,
                                double *errsoi_max, int32_t *index);
/* The patch part of EnergyBalanceCheck (BalanceCheckMod.F90:966-1003) on the
 * device-resident state after a full land step, for non-urban patches:
 *   errseb = sabv + sabg_chk + forc_lwrad - eflx_lwrad_out - eflx_sh_tot
 *            - eflx_lh_tot - eflx_soil_grnd - dhsdt_canopy
 *   errlon = eflx_lwrad_out - eflx_lwrad_net - forc_lwrad
 * (0 for inactive patches), written per filter entry to the device arrays
 * errseb / errlon when they are not NULL; waits and returns the largest
 * |errseb| and |errlon| with the lowest patch holding each.  The caller
 * applies the reference's thresholds (1e-7 W/m2 warning, 1e-5 W/m2 abort
 * after the skip steps).  errsol is ctsm_solar_balance_check_dev. */
int ctsm_energy_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                                  const ctsm_patch_state *pch, const ctsm_canopy_state *can,
                                  const ctsm_soilflux_state *sf, int32_t num_filterp, const int32_t *filterp_dev,
                                  double *errseb, double *errlon, double *errseb_max, int32_t *seb_patch,
                                  double *errlon_max, int32_t *lon_patch);

/* The solar part of EnergyBalanceCheck (BalanceCheckMod.F90:964-970, 1008-
 * 1012) for non-urban patches:
 *   errsol = fsa + fsr - (forc_solad_col(c,1) + forc_solad_col(c,2)
 *                         + forc_solai(g,1) + forc_solai(g,2))
 * with c = can->column(p), g = can->gridcell(p); 0 for inactive patches.
 * fsa / fsr are (p) arrays (SolarAbsorbedType / SurfaceAlbedoType), forc_solad_col
 * is (begc:endc, 1:2) and forc_solai (begg:endg, 1:2), all device arrays.
 * Writes errsol per filter entry when not NULL; waits and returns the largest
 * |errsol| and the lowest patch holding it (0 with no patches). */
int ctsm_solar_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_canopy_state *can,
                                 int32_t num_filterp, const int32_t *filterp_dev, const double *fsa,
                                 const double *fsr, const double *forc_solad_col, const double *forc_solai,
                                 double *errsol, double *errsol_max, int32_t *sol_patch);

/* Column water stores and fluxes of the water balance check.  The stores the
 * path itself holds come from ctsm_col_state (snl, patchi / patchf,
 * h2osoi_liq / h2osoi_ice [SNO], h2osfc, h2osno_no_layers) and ctsm_canopy_state (active, wtcol,
 * liqcan, snocan); these are the rest.  All (c) arrays, device memory. */
typedef struct ctsm_water_balance_state {
    const double  *total_plant_stored_h2o; /* waterdiagnostic%total_plant_stored_h2o_col (mm)  */
    const double  *wa;                     /* waterstate%wa_col: unconfined aquifer (mm)       */
    const double  *excess_ice;             /* (c, SOI) waterstate%excess_ice_col; NULL: zero   */
    const double  *dynbal_baseline_liq;    /* read when the baselines are subtracted           */
    const double  *dynbal_baseline_ice;
    const int32_t *hydrologically_active;  /* col%hydrologically_active                        */
    const int32_t *col_active;             /* col%active                                       */
    /* the step's fluxes (mm/s), BalanceCheckMod.F90:592-605 */
    const double  *forc_rain, *forc_snow, *qflx_floodc, *qflx_sfc_irrig, *qflx_glcice_dyn_water_flux;
    const double  *qflx_evap_tot, *qflx_surf, *qflx_qrgwl, *qflx_drain, *qflx_drain_perched;
    const double  *qflx_ice_runoff, *qflx_snwcp_discarded_liq, *qflx_snwcp_discarded_ice;
    const double  *begwb, *endwb;          /* column water mass at the step's start / end (mm)  */
    double        *errh2o;                 /* out: errh2o_col (mm); NULL: not written           */
} ctsm_water_balance_state;

/* ComputeWaterMassNonLake (TotalWaterAndHeatMod.F90:92-141, 200-393) for soil
 * and crop columns: per filter column, liquid = p2c(liqcan) +
 * total_plant_stored_h2o + the snow layers' liquid + (wa -
 * aquifer_water_baseline when hydrologically active) + h2osfc + the soil
 * layers' liquid (1..nlevgrnd); ice = p2c(snocan) + h2osno_no_layers + the
 * snow layers' ice + the soil layers' ice and excess ice; less the dynbal
 * baselines when subtract_dynbal_baselines != 0; water_mass = liquid + ice,
 * summed in the reference's order.  p2c is p2c_1d_filter (subgridAveMod.F90:
 * 292-320: patches patchi..patchf, active ones weighted by wtcol).  The
 * reference calls it for begwb (BalanceCheckMod.F90:426, baselines not
 * subtracted) and endwb (HydrologyDrainageMod.F90:182).  Enqueued on the
 * context's stream; water_mass is a device (c) array.  A filter entry outside
 * the layout, a patch range outside it or snl outside [-nlevsno, 0] is
 * reported (CTSM_E_ARG with the column) by the next call that waits. */
int ctsm_water_mass_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                        const ctsm_canopy_state *can, const ctsm_water_balance_state *wb, int32_t num_nolakec,
                        const int32_t *filter_nolakec_dev, double aquifer_water_baseline,
                        int32_t subtract_dynbal_baselines, double *water_mass);

/* The column water balance check of BalanceCheck (BalanceCheckMod.F90:588-
 * 630): per filter column
 *   errh2o = endwb - begwb - (forc_rain + forc_snow + qflx_floodc + qflx_sfc_irrig
 *            + qflx_glcice_dyn_water_flux - qflx_evap_tot - qflx_surf - qflx_qrgwl
 *            - qflx_drain - qflx_drain_perched - qflx_ice_runoff
 *            - qflx_snwcp_discarded_liq - qflx_snwcp_discarded_ice) * dtime
 * (0 for inactive columns), written to wb->errh2o when not NULL; waits and
 * returns the largest |errh2o| and the lowest column holding it.  The caller
 * applies the reference's thresholds (h2o_warning_thresh, 1e-5 mm abort after
 * the skip steps). */
int ctsm_water_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_water_balance_state *wb,
                                 int32_t num_filterc, const int32_t *filterc_dev, double dtime,
                                 double *errh2o_max, int32_t *errh2o_col);

/* One land step with canopy: CanopyFluxes over the patch filter, then per
 * column SoilTemperature, SetSoilWaterFractions, the transpiration sink and
 * SoilWater (clm_driver.F90:766, 900, 936; HydrologyNoDrainageMod.F90:302-346).
 * The patch arrays CanopyFluxes writes and SoilTemperature reads must be the
 * same device arrays in `pch` and `can`, and col->qflx_rootsoi must be
 * can->qflx_rootsoi_col. */
int ctsm_land_step_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                       const ctsm_col_state *col, const ctsm_patch_state *pch,
                       const ctsm_canopy_state *can,
                       int32_t num_filterc, const int32_t *filterc_dev,
                       int32_t num_filterp, const int32_t *filterp_dev, double dtime);

/* The land step with the reference's bookkeeping around it: t_ssbef and
 * t_h2osfc_bef saved first (BiogeophysPreFluxCalcsMod.F90:302-331), then
 * CanopyFluxes, SoilTemperature and SetSoilWaterFractions, SoilFluxes
 * (clm_driver.F90:921), the transpiration sink and, when sw is not NULL, the
 * column evaporation fluxes (clm_drv_patch2col) and the surface water input
 * chain that sets qflx_infl (HydrologyNoDrainageMod.F90:306-336), then
 * SoilWater and, when wt is not NULL, the column dew / sublimation fluxes
 * (clm_drv_patch2col), the water tables and RenewCondensation
 * (HydrologyNoDrainageMod.F90:355-373) and, when dr is not NULL (it needs
 * hprm), the drainage of HydrologyDrainage (clm_driver.F90:1148).  Arrays shared by two structs must be
 * the same device arrays, as for ctsm_land_step_dev, the CanopyFluxes outputs
 * SoilFluxes reads (sf->cgrnds, ...) must be the canopy state's,
 * sw->qflx_infl / sw->h2osfc must be col->qflx_infl / col->h2osfc, and
 * wt->zwt / zwt_perched / frost_table must be sw's. */
int ctsm_land_step_full_dev(ctsm_ctx *ctx, const ctsm_layout *lay,
                            const ctsm_col_state *col, const ctsm_patch_state *pch,
                            const ctsm_canopy_state *can, const ctsm_soilflux_state *sf,
                            const ctsm_surface_water_state *sw, const ctsm_hydro_params *hprm,
                            const ctsm_water_table_state *wt, const ctsm_drainage_state *dr,
                            int32_t num_filterc, const int32_t *filterc_dev,
                            int32_t num_filterp, const int32_t *filterp_dev, double dtime);

/* ---- captured step sequences (no counterpart in the reference: a batched
 * run replays the same launches every step, so they are captured once into a
 * HIP graph and replayed as one launch) -----------------------------------------
 * ctsm_graph_capture records `nsteps` consecutive steps over device-resident
 * state, with the context's current tuning: ctsm_land_step_dev when `can` is
 * not NULL, else ctsm_soil_step_fused_dev (pch is then the soil step's patch
 * state).  The graph holds the state pointers, filters and dtime given here
 * and its own workspace; nothing is launched by the capture.  Each
 * ctsm_graph_launch enqueues the steps on the context's stream, exactly the
 * kernels the eager calls launch; failures are reported by the next ctsm_sync
 * as for eager steps.  Destroy graphs before their context. */
typedef struct ctsm_graph ctsm_graph;
int ctsm_graph_capture(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_patch_state *pch, const ctsm_canopy_state *can,
                       int32_t num_filterc, const int32_t *filterc_dev,
                       int32_t num_filterp, const int32_t *filterp_dev, double dtime,
                       int32_t nsteps, ctsm_graph **out);
/* The same for `nsteps` full land steps (ctsm_land_step_full_dev with the
 * given SoilFluxes, surface water, water-table and drainage states; sw, hprm,
 * wt and dr may be NULL as there).  The balance check is not part of it: it
 * waits for the stream. */
int ctsm_graph_capture_full(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                            const ctsm_patch_state *pch, const ctsm_canopy_state *can,
                            const ctsm_soilflux_state *sf, const ctsm_surface_water_state *sw,
                            const ctsm_hydro_params *hprm, const ctsm_water_table_state *wt,
                            const ctsm_drainage_state *dr, int32_t num_filterc, const int32_t *filterc_dev,
                            int32_t num_filterp, const int32_t *filterp_dev, double dtime,
                            int32_t nsteps, ctsm_graph **out);
int ctsm_graph_launch(ctsm_ctx *ctx, ctsm_graph *graph);
int ctsm_graph_destroy(ctsm_graph *graph);

/* ---- history buffers on device (histFileMod.F90:1362-2169 hist_update_hbuf,
 * :2302-2358 hfields_normalize, :2361-2381 hfields_zero) --------------------------
 * A history field kept at the resolution of its source, as the reference's
 * branch that does not average up to the gridcell: n1d points x num2d levels of
 * `field`, level j (0-based) at field + j * ld.  `active` (n1d flags; NULL: every
 * point counts) is the patch / column `active` flag the reference checks.  hbuf
 * and nacs are num2d x n1d, point fastest, in device memory.  avgflag is the
 * flag's character: 'A' time average, 'S' sum, 'I' instantaneous, 'X' maximum,
 * 'M' minimum ('L', local solar time, is not taken).  Values equal to 1e36
 * (spval) are missing, as in the reference.  The descriptors are passed by
 * value to the kernels, so the calls may be captured in a graph. */
typedef struct ctsm_hist_field {
    const double  *field;
    const int32_t *active;
    double        *hbuf;
    int32_t       *nacs;
    int64_t        ld;
    int32_t        n1d;
    int32_t        num2d;
    int32_t        avgflag;
    int32_t        pad;
} ctsm_hist_field;

/* Accumulates every field into its buffer (hist_update_hbuf, one step). */
int ctsm_hist_update_dev(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t nfields);
/* Divides the 'A' buffers by their counts (hfields_normalize, before a write). */
int ctsm_hist_normalize_dev(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t nfields);
/* Clears buffers and counts (hfields_zero, after a write). */
int ctsm_hist_zero_dev(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t nfields);

/* ---- checkpoint / resume of device-resident state --------------------------
 * The drop-in path leaves restart files to the Fortran host (restFileMod.F90:
 * 74-156 writes the host arrays); a device-resident run writes its state here.
 * A checkpoint file holds named byte ranges: an 8-byte magic "CTSMCKP1", the
 * field count and, per field, a 64-byte name, its size, its file offset and a
 * 64-bit FNV-1a checksum of its 8-byte words, then the data, each range
 * aligned to 4 KiB.  Writing
 * waits for the context's stream, then streams the ranges through two pinned
 * staging buffers (the next chunk's copy overlaps the previous chunk's write)
 * into "<path>.tmp", flushes and fsyncs it and renames it over `path`: a
 * write that fails or is interrupted leaves an existing file at `path` as it
 * was.  Reading first hashes every named range in the file and returns
 * CTSM_E_IO, with the destination untouched, when a name is missing, a size
 * differs or a checksum does not match; only then does it fill the ranges
 * the same way.  `ptr` may be device or host memory. */
typedef struct ctsm_ckpt_field {
    const char *name;     /* NUL-terminated, at most 63 characters */
    void       *ptr;
    uint64_t    bytes;
} ctsm_ckpt_field;

int ctsm_checkpoint_write(ctsm_ctx *ctx, const char *path, const ctsm_ckpt_field *fields, int32_t nfields);
int ctsm_checkpoint_read(ctsm_ctx *ctx, const char *path, const ctsm_ckpt_field *fields, int32_t nfields);

/* ---- device memory and timing helpers --------------------------------------
 * Hosts need no other GPU runtime: device-resident state is allocated, filled
 * and read back through these (all on the context's device and stream). */
int ctsm_dev_alloc(ctsm_ctx *ctx, uint64_t bytes, void **out);
int ctsm_dev_free(ctsm_ctx *ctx, void *ptr);
/* kind: 1 = host->device, 2 = device->host, 3 = device->device.  Synchronous. */
int ctsm_memcpy(ctsm_ctx *ctx, void *dst, const void *src, uint64_t bytes, int32_t kind);
/* 2-D copy (rows of `width` bytes, `height` rows, pitches in bytes), e.g. one
 * member's column range of a level-major field; kind as ctsm_memcpy.  Synchronous. */
int ctsm_memcpy2d(ctsm_ctx *ctx, void *dst, uint64_t dpitch, const void *src, uint64_t spitch,
                  uint64_t width, uint64_t height, int32_t kind);
/* Device-to-device copy by a streaming kernel on the context's stream
 * (enqueues only; used for state snapshots). */
int ctsm_dev_copy(ctsm_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* Records an event on the context's stream (created on first use). */
int ctsm_event_record(ctsm_ctx *ctx, void **event);
/* Milliseconds between two recorded events (waits for the second). */
int ctsm_event_elapsed(void *ev_start, void *ev_stop, float *ms);
int ctsm_event_destroy(void *event);

/* Waits for the context's stream; returns CTSM_E_SOLVER if a column failed. */
int ctsm_sync(ctsm_ctx *ctx);

/* The context's HIP stream (hipStream_t), for callers that order their own work
 * after the library's, and for timing with events on that stream. */
void *ctsm_stream(ctsm_ctx *ctx);

/* ---- PFT rows of one member (no counterpart in the reference) -------------
 * Copies pft's npft rows into rows first_row .. first_row + npft - 1 of the
 * context's PFT table (after ctsm_set_canopy_params), waiting for the copy:
 * a perturbed-parameter member runs on its own rows. */
int ctsm_set_pft_rows(ctsm_ctx *ctx, int32_t first_row, const ctsm_pft_params *pft);

/* ---- ensemble server: continuous batching (no counterpart in the reference)
 * The reference runs one configuration per job; this serves many small
 * perturbed-parameter members ("requests") in one resident pool, the
 * per-clump call of clm_driver.F90:525 repeated over whichever members are
 * resident.  The pool is device state of ldc columns / ldp patches /
 * gridcells (one per column) that the caller allocated (ctsm_dev_alloc) and
 * describes once; the server hands out slots -- contiguous column, patch and
 * PFT-row ranges, first fit, so admission is by the pool's HBM budget --
 * and steps every running member together with full land steps
 * (ctsm_land_step_full_dev over the running columns and patches) each
 * followed by the energy balance checks (errsoi per column, errseb / errlon
 * per patch).  Data movement stays with the caller, who
 * knows its arrays:
 *   submit -> (progress admits it: state ADMITTED, slot in poll's col0 /
 *   pat0 / pft0) -> the caller copies the member's state into the slot on the
 *   context's stream, renumbering patchi / patchf / column / gridcell /
 *   itype by the slot offsets -> start -> (progress runs its steps: RUNNING)
 *   -> DONE once the last step has completed on the device -> the caller
 *   copies the state out -> release (frees the slot; also cancels a QUEUED
 *   or ADMITTED member).
 * progress never waits for the device unless wait != 0 (wait 1: for the
 * oldest tick in flight): it retires completed ticks (event queries), admits
 * queued members first-come while their ranges fit, and enqueues the next tick --
 * min(max_steps_per_tick, the fewest steps any running member still needs)
 * steps -- while fewer than two ticks are in flight, so admission and
 * retirement overlap the device's work.  Columns are independent: a member's
 * results do not depend on what shares its ticks.  wait 2 quiesces the
 * server: it waits for every tick in flight and retires them, and admits and
 * enqueues nothing (the pool then holds whole steps of every running member,
 * for a checkpoint).  Not thread-safe: one host thread drives a server. */
#define CTSM_SERVE_QUEUED    0
#define CTSM_SERVE_ADMITTED  1
#define CTSM_SERVE_RUNNING   2
#define CTSM_SERVE_DONE      3
#define CTSM_SERVE_FINISHING 4   /* internal: last step enqueued (poll reports RUNNING) */
typedef struct ctsm_server ctsm_server;
/* sw / hprm / wt / dr as for ctsm_land_step_full_dev (NULL leaves that part
 * out); members take PFT rows from pft_first on; dtime for every step. */
int ctsm_serve_create(ctsm_ctx *ctx, const ctsm_layout *pool, const ctsm_col_state *col,
                      const ctsm_patch_state *pch, const ctsm_canopy_state *can, const ctsm_soilflux_state *sf,
                      const ctsm_surface_water_state *sw, const ctsm_hydro_params *hprm,
                      const ctsm_water_table_state *wt, const ctsm_drainage_state *dr, int32_t pft_first,
                      int32_t max_steps_per_tick, double dtime, ctsm_server **out);
int ctsm_serve_submit(ctsm_server *s, int32_t ncol, int32_t npatch, int32_t npft, int32_t nsteps, int64_t *ticket);
int ctsm_serve_progress(ctsm_server *s, int32_t wait);
/* state (CTSM_SERVE_*), once admitted the slot -- 0-based offsets of the
 * member's first column / patch in the pool and its first PFT row -- and
 * the steps not yet enqueued (any pointer but state may be NULL). */
int ctsm_serve_poll(const ctsm_server *s, int64_t ticket, int32_t *state, int32_t *col0, int32_t *pat0,
                    int32_t *pft0, int32_t *left);
int ctsm_serve_start(ctsm_server *s, int64_t ticket);
int ctsm_serve_release(ctsm_server *s, int64_t ticket);
/* A RUNNING member at a given slot with `left` steps to go (a pool resumed
 * from a checkpoint). */
int ctsm_serve_resume_member(ctsm_server *s, int32_t ncol, int32_t npatch, int32_t npft, int32_t left,
                             int32_t col0, int32_t pat0, int32_t pft0, int64_t *ticket);
/* column-steps run, the largest |errsoi| seen (W/m2) and its pool column,
 * the largest |errseb| and |errlon| seen (W/m2) and their pool patches (0
 * when the state lacks a field the energy check reads), ticks in flight;
 * any pointer may be NULL */
int ctsm_serve_stats(const ctsm_server *s, int64_t *column_steps, double *worst_errsoi, int32_t *worst_col,
                     double *worst_errseb, int32_t *seb_patch, double *worst_errlon, int32_t *lon_patch,
                     int32_t *inflight);
int ctsm_serve_destroy(ctsm_server *s);

#ifdef __cplusplus
}
#endif

#endif /* CTSM_GPU_H */
