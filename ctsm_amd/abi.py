"""ctypes mirror of include/ctsm_gpu.h (the library's C-ABI).

The structs are declared once here and shared by the product loader
(ctsm_amd.lib) and the test-side oracle loader.  Field order must match the
header exactly; tests/test_abi.py checks sizes and offsets against the header
by compiling a probe with gcc.
"""
import ctypes as C

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class Config(C.Structure):
    _fields_ = [
        ("nlevsno", C.c_int32),
        ("nlevsoi", C.c_int32),
        ("nlevgrnd", C.c_int32),
        ("soilwater_movement_method", C.c_int32),
        ("upper_boundary_condition", C.c_int32),
        ("lower_boundary_condition", C.c_int32),
        ("flux_calculation", C.c_int32),
        ("snow_thermal_cond_method", C.c_int32),
        ("itmax_canopy_fluxes", C.c_int32),
        ("pad0", C.c_int32),
        ("dtmin", C.c_double),
        ("verySmall", C.c_double),
        ("xTolerUpper", C.c_double),
        ("xTolerLower", C.c_double),
        ("e_ice", C.c_double),
        ("z0param_method", C.c_int32),
        ("stomatalcond_method", C.c_int32),
        ("soil_resis_method", C.c_int32),
        ("use_biomass_heat_storage", C.c_int32),
        ("use_undercanopy_stability", C.c_int32),
        ("light_inhibit", C.c_int32),
        ("use_hydrstress", C.c_int32),
        ("use_excess_ice", C.c_int32),
        ("modifyphoto_and_lmr_forcrop", C.c_int32),
        ("pad1", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("richards_lanes", C.c_int32),
        ("richards_block", C.c_int32),
        ("soiltemp_block", C.c_int32),
        ("fused_single_kernel", C.c_int32),
        ("canopy_first_iterations", C.c_int32),
        ("canopy_round_iterations", C.c_int32),
        ("step_streams", C.c_int32),
        ("plain_routines", C.c_int32),
        ("canopy_waves", C.c_int32),
    ]


class Layout(C.Structure):
    _fields_ = [
        ("begc_alloc", C.c_int32),
        ("ldc", C.c_int32),
        ("begp_alloc", C.c_int32),
        ("ldp", C.c_int32),
        ("begg_alloc", C.c_int32),
        ("ldg", C.c_int32),
    ]


# (name, ctype, level-kind) in header order.  Level kinds:
#   c   (c)                 p   (p)
#   sno [-nlevsno+1 : nlevgrnd]   zi [-nlevsno : nlevgrnd]
#   soi [1 : nlevgrnd]      snow [-nlevsno+1 : 0]   lyr [-nlevsno+1 : 1] (patch)
COL_FIELDS = [
    ("snl", "i32", "c"),
    ("nbedrock", "i32", "c"),
    ("patchi", "i32", "c"),
    ("patchf", "i32", "c"),
    ("dz", "f64", "sno"),
    ("z", "f64", "sno"),
    ("zi", "f64", "zi"),
    ("watsat", "f64", "soi"),
    ("bsw", "f64", "soi"),
    ("sucsat", "f64", "soi"),
    ("hksat", "f64", "soi"),
    ("tkmg", "f64", "soi"),
    ("tkdry", "f64", "soi"),
    ("csol", "f64", "soi"),
    ("eff_porosity", "f64", "soi"),
    ("smp_l", "f64", "soi"),
    ("hk_l", "f64", "soi"),
    ("thk", "f64", "sno"),
    ("icefrac", "f64", "soi"),
    ("qcharge", "f64", "c"),
    ("num_substeps", "f64", "c"),
    ("qin", "f64", "soi"),
    ("qout", "f64", "soi"),
    ("h2osoi_liq", "f64", "sno"),
    ("h2osoi_ice", "f64", "sno"),
    ("h2osfc", "f64", "c"),
    ("h2osno_no_layers", "f64", "c"),
    ("int_snow", "f64", "c"),
    ("frac_sno_eff", "f64", "c"),
    ("frac_h2osfc", "f64", "c"),
    ("snow_depth", "f64", "c"),
    ("bw", "f64", "snow"),
    ("snomelt_accum", "f64", "c"),
    ("qflx_infl", "f64", "c"),
    ("qflx_rootsoi", "f64", "soi"),
    ("qflx_snomelt", "f64", "c"),
    ("qflx_snofrz", "f64", "c"),
    ("qflx_snow_drain", "f64", "c"),
    ("qflx_snomelt_lyr", "f64", "snow"),
    ("qflx_snofrz_lyr", "f64", "snow"),
    ("qflx_h2osfc_to_ice", "f64", "c"),
    ("t_soisno", "f64", "sno"),
    ("t_h2osfc", "f64", "c"),
    ("t_grnd", "f64", "c"),
    ("fact", "f64", "sno"),
    ("xmf", "f64", "c"),
    ("xmf_h2osfc", "f64", "c"),
    ("c_h2osfc", "f64", "c"),
    ("imelt", "i32", "sno"),
    ("emg", "f64", "c"),
    ("htvp", "f64", "c"),
    ("eflx_bot", "f64", "c"),
    ("eflx_fgr12", "f64", "c"),
    ("eflx_fgr", "f64", "soi"),
    ("eflx_snomelt", "f64", "c"),
    ("eflx_h2osfc_to_snow", "f64", "c"),
    ("forc_lwrad", "f64", "c"),
    ("zwt", "f64", "c"),
    ("smpmin", "f64", "c"),
]

PATCH_FIELDS = [
    ("column", "i32", "p"),
    ("wtcol", "f64", "p"),
    ("frac_veg_nosno", "i32", "p"),
    ("sabg", "f64", "p"),
    ("sabg_soil", "f64", "p"),
    ("sabg_snow", "f64", "p"),
    ("sabg_lyr", "f64", "lyr"),
    ("dlrad", "f64", "p"),
    ("cgrnd", "f64", "p"),
    ("eflx_sh_grnd", "f64", "p"),
    ("eflx_sh_snow", "f64", "p"),
    ("eflx_sh_soil", "f64", "p"),
    ("eflx_sh_h2osfc", "f64", "p"),
    ("qflx_evap_soi", "f64", "p"),
    ("qflx_ev_snow", "f64", "p"),
    ("qflx_ev_soil", "f64", "p"),
    ("qflx_ev_h2osfc", "f64", "p"),
    ("eflx_gnet", "f64", "p"),
    ("dgnetdT", "f64", "p"),
    ("sabg_chk", "f64", "p"),
]

# CanopyFluxes state; extra level kinds: g (gridcell), soip (patch x nlevgrnd, lo 1)
CANOPY_FIELDS = [
    ("itype", "i32", "p"),
    ("column", "i32", "p"),
    ("gridcell", "i32", "p"),
    ("wtcol", "f64", "p"),
    ("frac_veg_nosno", "i32", "p"),
    ("active", "i32", "p"),
    ("forc_u", "f64", "g"),
    ("forc_v", "f64", "g"),
    ("forc_pco2", "f64", "g"),
    ("forc_po2", "f64", "g"),
    ("forc_hgt_u", "f64", "g"),
    ("forc_hgt_t", "f64", "g"),
    ("forc_hgt_q", "f64", "g"),
    ("dayl", "f64", "g"),
    ("max_dayl", "f64", "g"),
    ("near_local_noon", "i32", "g"),
    ("forc_q", "f64", "c"),
    ("forc_pbot", "f64", "c"),
    ("forc_th", "f64", "c"),
    ("forc_rho", "f64", "c"),
    ("thv", "f64", "c"),
    ("qg", "f64", "c"),
    ("qg_snow", "f64", "c"),
    ("qg_soil", "f64", "c"),
    ("qg_h2osfc", "f64", "c"),
    ("dqgdT", "f64", "c"),
    ("z0mg", "f64", "c"),
    ("soilbeta", "f64", "c"),
    ("soilresis", "f64", "c"),
    ("h2osoi_liqvol", "f64", "soi"),
    ("qflx_tran_veg_col", "f64", "c"),
    ("rootr_col", "f64", "soi"),
    ("qflx_rootsoi_col", "f64", "soi"),
    ("elai", "f64", "p"),
    ("esai", "f64", "p"),
    ("laisun", "f64", "p"),
    ("laisha", "f64", "p"),
    ("htop", "f64", "p"),
    ("sabv", "f64", "p"),
    ("parsun_z", "f64", "p"),
    ("parsha_z", "f64", "p"),
    ("laisun_z", "f64", "p"),
    ("laisha_z", "f64", "p"),
    ("vcmaxcintsun", "f64", "p"),
    ("vcmaxcintsha", "f64", "p"),
    ("nrad", "i32", "p"),
    ("thm", "f64", "p"),
    ("emv", "f64", "p"),
    ("fwet", "f64", "p"),
    ("fdry", "f64", "p"),
    ("t_a10", "f64", "p"),
    ("o3coefvsun", "f64", "p"),
    ("o3coefvsha", "f64", "p"),
    ("o3coefgsun", "f64", "p"),
    ("o3coefgsha", "f64", "p"),
    ("rootfr", "f64", "soip"),
    ("displa", "f64", "p"),
    ("z0mv", "f64", "p"),
    ("t_veg", "f64", "p"),
    ("t_stem", "f64", "p"),
    ("snocan", "f64", "p"),
    ("liqcan", "f64", "p"),
    ("cgrnds", "f64", "p"),
    ("cgrndl", "f64", "p"),
    ("z0hv", "f64", "p"),
    ("z0qv", "f64", "p"),
    ("forc_hgt_u_patch", "f64", "p"),
    ("forc_hgt_t_patch", "f64", "p"),
    ("forc_hgt_q_patch", "f64", "p"),
    ("ram1", "f64", "p"),
    ("rb1", "f64", "p"),
    ("rah1", "f64", "p"),
    ("rah2", "f64", "p"),
    ("raw1", "f64", "p"),
    ("raw2", "f64", "p"),
    ("ustar", "f64", "p"),
    ("um", "f64", "p"),
    ("uaf", "f64", "p"),
    ("taf", "f64", "p"),
    ("qaf", "f64", "p"),
    ("obu", "f64", "p"),
    ("zeta", "f64", "p"),
    ("vpd", "f64", "p"),
    ("num_iter", "f64", "p"),
    ("u10", "f64", "p"),
    ("u10_clm", "f64", "p"),
    ("va", "f64", "p"),
    ("fv", "f64", "p"),
    ("vds", "f64", "p"),
    ("btran", "f64", "p"),
    ("rresis", "f64", "soip"),
    ("rootr", "f64", "soip"),
    ("eflx_sh_veg", "f64", "p"),
    ("eflx_sh_stem", "f64", "p"),
    ("eflx_sh_grnd", "f64", "p"),
    ("eflx_sh_snow", "f64", "p"),
    ("eflx_sh_soil", "f64", "p"),
    ("eflx_sh_h2osfc", "f64", "p"),
    ("taux", "f64", "p"),
    ("tauy", "f64", "p"),
    ("dlrad", "f64", "p"),
    ("ulrad", "f64", "p"),
    ("cgrnd", "f64", "p"),
    ("dhsdt_canopy", "f64", "p"),
    ("qflx_evap_soi", "f64", "p"),
    ("qflx_ev_snow", "f64", "p"),
    ("qflx_ev_soil", "f64", "p"),
    ("qflx_ev_h2osfc", "f64", "p"),
    ("qflx_tran_veg", "f64", "p"),
    ("qflx_evap_veg", "f64", "p"),
    ("rh_af", "f64", "p"),
    ("q_ref2m", "f64", "p"),
    ("rh_ref2m", "f64", "p"),
    ("rh_ref2m_r", "f64", "p"),
    ("vpd_ref2m", "f64", "p"),
    ("iwue_ln", "f64", "p"),
    ("t_ref2m", "f64", "p"),
    ("t_ref2m_r", "f64", "p"),
    ("t_skin", "f64", "p"),
    ("stem_biomass", "f64", "p"),
    ("leaf_biomass", "f64", "p"),
    ("dleaf_patch", "f64", "p"),
    ("rssun", "f64", "p"),
    ("rssha", "f64", "p"),
    ("psnsun", "f64", "p"),
    ("psnsha", "f64", "p"),
    ("psnsun_wc", "f64", "p"),
    ("psnsun_wj", "f64", "p"),
    ("psnsun_wp", "f64", "p"),
    ("psnsha_wc", "f64", "p"),
    ("psnsha_wj", "f64", "p"),
    ("psnsha_wp", "f64", "p"),
    ("fpsn", "f64", "p"),
    ("fpsn_wc", "f64", "p"),
    ("fpsn_wj", "f64", "p"),
    ("fpsn_wp", "f64", "p"),
    ("lmrsun", "f64", "p"),
    ("lmrsha", "f64", "p"),
    ("gs_mol_sun", "f64", "p"),
    ("gs_mol_sha", "f64", "p"),
    ("gs_mol_sun_ln", "f64", "p"),
    ("gs_mol_sha_ln", "f64", "p"),
    ("cisun_z", "f64", "p"),
    ("cisha_z", "f64", "p"),
    ("rssun_z", "f64", "p"),
    ("rssha_z", "f64", "p"),
    ("lmrsun_z", "f64", "p"),
    ("lmrsha_z", "f64", "p"),
    ("psnsun_z", "f64", "p"),
    ("psnsha_z", "f64", "p"),
    ("vcmax_z", "f64", "p"),
    ("tpu_z", "f64", "p"),
    ("kp_z", "f64", "p"),
    ("ac", "f64", "p"),
    ("aj", "f64", "p"),
    ("ap", "f64", "p"),
    ("ag", "f64", "p"),
    ("an", "f64", "p"),
    ("gb_mol", "f64", "p"),
    ("gs_mol", "f64", "p"),
    ("cp", "f64", "p"),
    ("kc", "f64", "p"),
    ("ko", "f64", "p"),
    ("qe", "f64", "p"),
    ("bbb", "f64", "p"),
    ("mbb", "f64", "p"),
    ("rh_leaf", "f64", "p"),
    ("vpd_can", "f64", "p"),
    ("lnca", "f64", "p"),
    ("c3flag", "i32", "p"),
    ("canopy_err", "f64", "p"),
]

# ctsm_soilflux_state, in declaration order (kind: "p" patch, "c" column, "sno" snow+soil levels)
SOILFLUX_FIELDS = [
    ("active", "i32", "p"),
    ("t_ssbef", "f64", "sno"), ("t_h2osfc_bef", "f64", "c"),
    ("eflx_sh_grnd", "f64", "p"), ("qflx_evap_soi", "f64", "p"), ("qflx_ev_snow", "f64", "p"),
    ("qflx_ev_soil", "f64", "p"), ("qflx_ev_h2osfc", "f64", "p"),
    ("cgrnds", "f64", "p"), ("cgrndl", "f64", "p"), ("ulrad", "f64", "p"), ("eflx_sh_veg", "f64", "p"),
    ("eflx_sh_stem", "f64", "p"), ("qflx_evap_veg", "f64", "p"), ("qflx_tran_veg", "f64", "p"),
    ("eflx_soil_grnd", "f64", "p"), ("eflx_soil_grnd_r", "f64", "p"), ("eflx_sh_tot", "f64", "p"),
    ("eflx_sh_tot_r", "f64", "p"),
    ("qflx_evap_tot", "f64", "p"), ("eflx_lh_tot", "f64", "p"), ("eflx_lh_tot_r", "f64", "p"),
    ("qflx_evap_can", "f64", "p"), ("eflx_lh_vege", "f64", "p"), ("eflx_lh_vegt", "f64", "p"),
    ("eflx_lh_grnd", "f64", "p"),
    ("eflx_lwrad_out", "f64", "p"), ("eflx_lwrad_out_r", "f64", "p"), ("eflx_lwrad_net", "f64", "p"),
    ("eflx_lwrad_net_r", "f64", "p"), ("t_skin", "f64", "p"),
    ("qflx_liqevap_from_top_layer", "f64", "p"), ("qflx_solidevap_from_top_layer", "f64", "p"),
    ("qflx_liqdew_to_top_layer", "f64", "p"), ("qflx_soliddew_to_top_layer", "f64", "p"),
    ("errsoi_patch", "f64", "p"), ("errsoi_col", "f64", "c"),
]

# ctsm_surface_water_state, in declaration order (all column arrays)
SURFWATER_FIELDS = [(n, "f64", "c") for n in (
    "qflx_rain_plus_snomelt", "qflx_snow_h2osfc", "qflx_floodc",
    "qflx_liqevap_from_top_layer", "qflx_ev_soil", "qflx_ev_h2osfc",
    "zwt", "zwt_perched", "frost_table", "wtfact", "topo_slope", "h2osfc_thresh", "frac_h2osfc_nosnow",
    "h2osfc",
    "fsat", "fcov", "qflx_sat_excess_surf", "qflx_top_soil", "qflx_in_soil", "qflx_top_soil_to_h2osfc",
    "qinmax", "qflx_infl_excess", "qflx_in_soil_limited", "qflx_in_h2osfc", "qflx_infl_excess_surf",
    "qflx_h2osfc_surf", "qflx_h2osfc_drain", "qflx_infl", "qflx_surf")]


WATERTABLE_FIELDS = [("zwt", "f64", "c"), ("zwt_perched", "f64", "c"), ("frost_table", "f64", "c"),
                     ("h2osoi_vol", "f64", "soi"), ("qflx_liqdew_to_top_layer", "f64", "c"),
                     ("qflx_soliddew_to_top_layer", "f64", "c"), ("qflx_solidevap_from_top_layer", "f64", "c")]


DRAINAGE_FIELDS = [(n, "f64", "soi" if n == "h2osoi_vol" else "c") for n in (
    "topo_slope", "zwt_perched", "frost_table", "qflx_surf", "qflx_snwcp_liq", "zwt", "h2osoi_vol",
    "qflx_drain_perched", "qflx_latflow_out", "qflx_latflow_in", "qflx_rsub_sat", "qflx_ice_runoff_xs",
    "qflx_drain", "qflx_qrgwl", "qflx_runoff")]


# column water stores and fluxes of the water balance check (ctsm_water_balance_state)
WATERBAL_FIELDS = [
    ("total_plant_stored_h2o", "f64", "c"), ("wa", "f64", "c"),
    ("excess_ice", "f64", "soi"), ("dynbal_baseline_liq", "f64", "c"), ("dynbal_baseline_ice", "f64", "c"),
    ("hydrologically_active", "i32", "c"), ("col_active", "i32", "c")] + [(n, "f64", "c") for n in (
        "forc_rain", "forc_snow", "qflx_floodc", "qflx_sfc_irrig", "qflx_glcice_dyn_water_flux", "qflx_evap_tot",
        "qflx_surf", "qflx_qrgwl", "qflx_drain", "qflx_drain_perched", "qflx_ice_runoff",
        "qflx_snwcp_discarded_liq", "qflx_snwcp_discarded_ice", "begwb", "endwb", "errh2o")]

# plant hydraulic stress, soil side (ctsm_phs_state): patch members, then column members
PHS_FIELDS = [
    ("itype", "i32", "p"), ("column", "i32", "p"), ("active", "i32", "p"), ("frac_veg_nosno", "i32", "p"),
    ("wtcol", "f64", "p"), ("froot_carbon", "f64", "p"), ("tsai", "f64", "p"), ("tlai", "f64", "p"),
    ("rootfr", "f64", "soip"), ("vegwp", "f64", "veg"), ("k_soil_root", "f64", "soip"),
    ("root_conductance", "f64", "soip"), ("soil_conductance", "f64", "soip"), ("qflx_hydr_redist", "f64", "p"),
    ("patchi", "i32", "c"), ("patchf", "i32", "c"), ("dz", "f64", "sno"), ("z", "f64", "sno"),
    ("hksat", "f64", "soi"), ("hk_l", "f64", "soi"), ("smp_l", "f64", "soi"), ("qflx_rootsoi", "f64", "soi"),
    ("qflx_phs_neg", "f64", "c"),
]


class HydroParams(C.Structure):
    _fields_ = [("fff", C.c_double), ("pc", C.c_double), ("mu", C.c_double), ("aq_sp_yield_min", C.c_double),
                ("n_baseflow", C.c_double), ("baseflow_scalar", C.c_double), ("perched_baseflow_scalar", C.c_double),
                ("h2osfcflag", C.c_int32), ("pad", C.c_int32)]



CANOPY_PARAM_NAMES = [
    "cv", "a_coef", "a_exp", "csoilc", "z_dl", "lai_dl", "wind_min", "zetamaxstable",
    "kc25_coef", "ko25_coef", "cp25_yr2000", "kcha", "koha", "cpha",
    "vcmaxha", "jmaxha", "tpuha", "lmrha", "vcmaxhd", "jmaxhd", "tpuhd", "lmrhd", "lmrse",
    "fnr", "act25", "theta_psii", "theta_ip", "fnps", "tpu25ratio", "kp25ratio",
    "vcmaxse_sf", "jmaxse_sf", "tpuse_sf", "jmax25top_sf", "leaf_mr_vcm",
]


class CanopyParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in CANOPY_PARAM_NAMES]


PFT_FIELDS = [
    ("dleaf", "f64"), ("c3psn", "f64"), ("slatop", "f64"), ("leafcn", "f64"), ("flnr", "f64"), ("fnitr", "f64"),
    ("mbbopt", "f64"), ("medlynslope", "f64"), ("medlynintercept", "f64"),
    ("smpso", "f64"), ("smpsc", "f64"), ("theta_cj", "f64"),
    ("z0v_Cr", "f64"), ("z0v_Cs", "f64"), ("z0v_c", "f64"), ("z0v_cw", "f64"), ("z0v_LAImax", "f64"),
    ("is_tree", "i32"), ("is_shrub", "i32"),
    ("dbh", "f64"), ("fbw", "f64"), ("nstem", "f64"), ("rstem_per_dbh", "f64"), ("wood_density", "f64"),
]

_CT = {"i32": i32p, "f64": f64p}


class PftParams(C.Structure):
    _fields_ = [("npft", C.c_int32), ("pad", C.c_int32)] + [(n, _CT[t]) for n, t in PFT_FIELDS]


class CanopyState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in CANOPY_FIELDS]


class SurfaceWaterState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in SURFWATER_FIELDS]


class DrainageState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in DRAINAGE_FIELDS]


class WaterBalanceState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in WATERBAL_FIELDS]


class PhsState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in PHS_FIELDS]


class PhsParams(C.Structure):
    _fields_ = [("npft", C.c_int32), ("pad", C.c_int32)] + [
        (n, f64p) for n in ("krmax", "root_radius", "root_density", "froot_leaf", "psi50", "ck", "kmax", "crop")]


# CanopyFluxes with plant hydraulic stress (ctsm_phs_canopy_state); phs2 =
# (begp:endp, sun / sha)
PHS_CANOPY_FIELDS = [
    ("tlai_z", "f64", "p"), ("local_morning", "i32", "p"), ("vegwp", "f64", "veg"), ("vegwp_ln", "f64", "veg"),
    ("vegwp_pd", "f64", "veg"), ("bsun", "f64", "p"), ("bsha", "f64", "p"), ("an_sun", "f64", "p"),
    ("an_sha", "f64", "p"), ("vcmax_z_phs", "f64", "phs2"), ("tpu_z_phs", "f64", "phs2"), ("kp_z_phs", "f64", "phs2"),
    ("ac_phs", "f64", "phs2"), ("aj_phs", "f64", "phs2"), ("ap_phs", "f64", "phs2"), ("ag_phs", "f64", "phs2"),
]


class PhsCanopyState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in PHS_CANOPY_FIELDS]


# calcstress (ctsm_phs_stress_state): patch members, then column members
PHS_STRESS_FIELDS = [
    ("itype", "i32", "p"), ("column", "i32", "p"), ("local_morning", "i32", "p"),
    ("laisun", "f64", "p"), ("laisha", "f64", "p"), ("elai", "f64", "p"), ("esai", "f64", "p"),
    ("tsai", "f64", "p"), ("htop", "f64", "p"), ("fdry", "f64", "p"), ("thm", "f64", "p"),
    ("gb_mol", "f64", "p"), ("gs_mol_sun", "f64", "p"), ("gs_mol_sha", "f64", "p"), ("qsatl", "f64", "p"),
    ("qaf", "f64", "p"), ("k_soil_root", "f64", "soip"), ("vegwp", "f64", "veg"), ("vegwp_pd", "f64", "veg"),
    ("bsun", "f64", "p"), ("bsha", "f64", "p"), ("qflx_tran_veg", "f64", "p"),
    ("forc_rho", "f64", "c"), ("forc_pbot", "f64", "c"), ("z", "f64", "sno"), ("smp_l", "f64", "soi"),
]


class PhsStressState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in PHS_STRESS_FIELDS]


class WaterTableState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in WATERTABLE_FIELDS]


class HistField(C.Structure):
    _fields_ = [("field", C.c_void_p), ("active", C.c_void_p), ("hbuf", C.c_void_p), ("nacs", C.c_void_p),
                ("ld", C.c_int64), ("n1d", C.c_int32), ("num2d", C.c_int32), ("avgflag", C.c_int32),
                ("pad", C.c_int32)]


class CkptField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ptr", C.c_void_p), ("bytes", C.c_uint64)]


class SoilFluxState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in SOILFLUX_FIELDS]


class ColState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in COL_FIELDS]


class PatchState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in PATCH_FIELDS]


def nlev_of(kind, nlevsno, nlevgrnd):
    """Number of levels stored for a level kind (1 for scalars)."""
    return {
        "c": 1,
        "p": 1,
        "sno": nlevsno + nlevgrnd,
        "zi": nlevsno + nlevgrnd + 1,
        "soi": nlevgrnd,
        "snow": nlevsno,
        "lyr": nlevsno + 1,
        "g": 1,
        "soip": nlevgrnd,
    }[kind]


def lo_of(kind, nlevsno):
    """Fortran lower level bound of a level kind."""
    return {"c": 0, "p": 0, "g": 0, "sno": 1 - nlevsno, "zi": -nlevsno, "soi": 1, "soip": 1,
            "snow": 1 - nlevsno, "lyr": 1 - nlevsno}[kind]
