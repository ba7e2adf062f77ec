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

/* soil water lower / upper boundary conditions (SoilWaterMovementMod.F90:55-59) */
#define CTSM_BC_HEAD        0
#define CTSM_BC_FLUX        1
#define CTSM_BC_ZERO_FLUX   2
#define CTSM_BC_WATERTABLE  3
#define CTSM_BC_AQUIFER     4

/* snow thermal conductivity (SoilTemperatureMod.F90:739-760) */
#define CTSM_SNOWTHK_JORDAN1991 0
#define CTSM_SNOWTHK_STURM1997  1

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
    int32_t lower_boundary_condition;  /* CTSM_BC_ZERO_FLUX or CTSM_BC_FLUX       */
    int32_t flux_calculation;   /* 1 = inexpensive, 42 = expensive                  */
    int32_t snow_thermal_cond_method;  /* CTSM_SNOWTHK_*                          */
    int32_t itmax_canopy_fluxes;/* 40                                               */
    int32_t pad0;
    double  dtmin;              /* 60 s                                             */
    double  verySmall;          /* 1e-8                                             */
    double  xTolerUpper;        /* 0.1                                              */
    double  xTolerLower;        /* 0.01                                             */
    double  e_ice;              /* 6 (ice impedance exponent)                       */
} ctsm_config;

/* ---- array layout of one call -------------------------------------------- */
typedef struct ctsm_layout {
    int32_t begc_alloc;   /* lower column bound of the allocation (bounds_proc%begc) */
    int32_t ldc;          /* column extent of the allocation                          */
    int32_t begp_alloc;   /* lower patch bound of the allocation                      */
    int32_t ldp;          /* patch extent of the allocation                           */
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

/* ---- launch shapes (no counterpart in the reference: set by the autotuner) --
 * richards_lanes: lanes cooperating on one column in the Richards solve
 *   (1, 2, 4 or 8); the heat solve always runs one column per lane.
 * fused_single_kernel: 1 = the whole soil step in one kernel (one lane per
 *   column), 0 = heat solve kernel + Richards kernel (2 launches per step). */
typedef struct ctsm_tuning {
    int32_t richards_lanes;
    int32_t richards_block;
    int32_t soiltemp_block;
    int32_t fused_single_kernel;
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

/* ---- device memory and timing helpers --------------------------------------
 * Hosts need no other GPU runtime: device-resident state is allocated, filled
 * and read back through these (all on the context's device and stream). */
int ctsm_dev_alloc(ctsm_ctx *ctx, uint64_t bytes, void **out);
int ctsm_dev_free(ctsm_ctx *ctx, void *ptr);
/* kind: 1 = host->device, 2 = device->host, 3 = device->device.  Synchronous. */
int ctsm_memcpy(ctsm_ctx *ctx, void *dst, const void *src, uint64_t bytes, int32_t kind);
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

#ifdef __cplusplus
}
#endif

#endif /* CTSM_GPU_H */
