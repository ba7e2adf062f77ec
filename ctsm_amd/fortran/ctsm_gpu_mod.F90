module ctsm_gpu_mod
  !-----------------------------------------------------------------------
  ! ISO_C_BINDING interface of libctsm_mi355x.so (include/ctsm_gpu.h).
  !
  ! The derived types below are interoperable mirrors of the C structs; the
  ! state structs hold c_ptr members that the caller sets with c_loc() of
  ! the reference's own arrays (e.g. c_loc(waterstatebulk_inst%h2osoi_liq_col)),
  ! so the library reads and writes CTSM's memory in place.
  !
  ! A drop-in wrapper for SoilWater (SoilWaterMovementMod.F90:240-243) or
  ! SoilTemperature (SoilTemperatureMod.F90:92-95) built on this module fills
  ! the structs from the *_inst objects, converts nothing (filters stay
  ! 1-based reference indices) and calls endrun with the failing column when
  ! a call returns CTSM_E_SOLVER.
  !-----------------------------------------------------------------------
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CTSM_OK = 0, CTSM_E_ARG = 1, CTSM_E_CONFIG = 2, &
       CTSM_E_HIP = 3, CTSM_E_SOLVER = 4, CTSM_E_NODEVICE = 5, CTSM_E_ENDRUN = 6, CTSM_E_IO = 7
  integer(c_int), parameter, public :: CTSM_Z0_ZENGWANG2007 = 0, CTSM_Z0_MEIER2022 = 1
  integer(c_int), parameter, public :: CTSM_STOMATA_BB1987 = 1, CTSM_STOMATA_MEDLYN2011 = 2
  integer(c_int), parameter, public :: CTSM_SOILRESIS_LEEPIELKE1992 = 0, CTSM_SOILRESIS_SL14 = 1
  integer(c_int), parameter, public :: CTSM_BC_FLUX = 1, CTSM_BC_ZERO_FLUX = 2
  integer(c_int), parameter, public :: CTSM_SNOWTHK_JORDAN1991 = 0, CTSM_SNOWTHK_STURM1997 = 1

  type, bind(C), public :: ctsm_config_t
     integer(c_int32_t) :: nlevsno, nlevsoi, nlevgrnd
     integer(c_int32_t) :: soilwater_movement_method
     integer(c_int32_t) :: upper_boundary_condition, lower_boundary_condition
     integer(c_int32_t) :: flux_calculation, snow_thermal_cond_method
     integer(c_int32_t) :: itmax_canopy_fluxes, pad0
     real(c_double)     :: dtmin, verySmall, xTolerUpper, xTolerLower, e_ice
     integer(c_int32_t) :: z0param_method, stomatalcond_method, soil_resis_method
     integer(c_int32_t) :: use_biomass_heat_storage, use_undercanopy_stability, light_inhibit
     integer(c_int32_t) :: use_hydrstress, use_excess_ice
     integer(c_int32_t) :: modifyphoto_and_lmr_forcrop, pad1
  end type ctsm_config_t

  type, bind(C), public :: ctsm_layout_t
     integer(c_int32_t) :: begc_alloc, ldc, begp_alloc, ldp
     integer(c_int32_t) :: begg_alloc = 1, ldg = 0
  end type ctsm_layout_t

  type, bind(C), public :: ctsm_col_state_t
     type(c_ptr) :: snl = c_null_ptr, nbedrock = c_null_ptr, patchi = c_null_ptr, patchf = c_null_ptr
     type(c_ptr) :: dz = c_null_ptr, z = c_null_ptr, zi = c_null_ptr
     type(c_ptr) :: watsat = c_null_ptr, bsw = c_null_ptr, sucsat = c_null_ptr, hksat = c_null_ptr
     type(c_ptr) :: tkmg = c_null_ptr, tkdry = c_null_ptr, csol = c_null_ptr
     type(c_ptr) :: eff_porosity = c_null_ptr, smp_l = c_null_ptr, hk_l = c_null_ptr, thk = c_null_ptr
     type(c_ptr) :: icefrac = c_null_ptr, qcharge = c_null_ptr, num_substeps = c_null_ptr
     type(c_ptr) :: qin = c_null_ptr, qout = c_null_ptr
     type(c_ptr) :: h2osoi_liq = c_null_ptr, h2osoi_ice = c_null_ptr, h2osfc = c_null_ptr
     type(c_ptr) :: h2osno_no_layers = c_null_ptr, int_snow = c_null_ptr
     type(c_ptr) :: frac_sno_eff = c_null_ptr, frac_h2osfc = c_null_ptr, snow_depth = c_null_ptr
     type(c_ptr) :: bw = c_null_ptr, snomelt_accum = c_null_ptr
     type(c_ptr) :: qflx_infl = c_null_ptr, qflx_rootsoi = c_null_ptr, qflx_snomelt = c_null_ptr
     type(c_ptr) :: qflx_snofrz = c_null_ptr, qflx_snow_drain = c_null_ptr
     type(c_ptr) :: qflx_snomelt_lyr = c_null_ptr, qflx_snofrz_lyr = c_null_ptr
     type(c_ptr) :: qflx_h2osfc_to_ice = c_null_ptr
     type(c_ptr) :: t_soisno = c_null_ptr, t_h2osfc = c_null_ptr, t_grnd = c_null_ptr, fact = c_null_ptr
     type(c_ptr) :: xmf = c_null_ptr, xmf_h2osfc = c_null_ptr, c_h2osfc = c_null_ptr, imelt = c_null_ptr
     type(c_ptr) :: emg = c_null_ptr, htvp = c_null_ptr, eflx_bot = c_null_ptr, eflx_fgr12 = c_null_ptr
     type(c_ptr) :: eflx_fgr = c_null_ptr, eflx_snomelt = c_null_ptr, eflx_h2osfc_to_snow = c_null_ptr
     type(c_ptr) :: forc_lwrad = c_null_ptr
     type(c_ptr) :: zwt = c_null_ptr, smpmin = c_null_ptr
  end type ctsm_col_state_t

  type, bind(C), public :: ctsm_patch_state_t
     type(c_ptr) :: column = c_null_ptr, wtcol = c_null_ptr, frac_veg_nosno = c_null_ptr
     type(c_ptr) :: sabg = c_null_ptr, sabg_soil = c_null_ptr, sabg_snow = c_null_ptr, sabg_lyr = c_null_ptr
     type(c_ptr) :: dlrad = c_null_ptr, cgrnd = c_null_ptr, eflx_sh_grnd = c_null_ptr
     type(c_ptr) :: eflx_sh_snow = c_null_ptr, eflx_sh_soil = c_null_ptr, eflx_sh_h2osfc = c_null_ptr
     type(c_ptr) :: qflx_evap_soi = c_null_ptr, qflx_ev_snow = c_null_ptr, qflx_ev_soil = c_null_ptr
     type(c_ptr) :: qflx_ev_h2osfc = c_null_ptr, eflx_gnet = c_null_ptr, dgnetdT = c_null_ptr
     type(c_ptr) :: sabg_chk = c_null_ptr
  end type ctsm_patch_state_t

  ! param-file scalars of CanopyFluxes / Photosynthesis / FrictionVelocity
  type, bind(C), public :: ctsm_canopy_params_t
     real(c_double) :: cv, a_coef, a_exp, csoilc, z_dl, lai_dl
     real(c_double) :: wind_min, zetamaxstable, kc25_coef, ko25_coef, cp25_yr2000, kcha
     real(c_double) :: koha, cpha, vcmaxha, jmaxha, tpuha, lmrha
     real(c_double) :: vcmaxhd, jmaxhd, tpuhd, lmrhd, lmrse, fnr
     real(c_double) :: act25, theta_psii, theta_ip, fnps, tpu25ratio, kp25ratio
     real(c_double) :: vcmaxse_sf, jmaxse_sf, tpuse_sf, jmax25top_sf, leaf_mr_vcm
  end type ctsm_canopy_params_t

  ! pftcon arrays (c_loc(pftcon%dleaf) etc., 0:mxpft)
  type, bind(C), public :: ctsm_pft_params_t
     integer(c_int32_t) :: npft = 0, pad = 0
     type(c_ptr) :: dleaf = c_null_ptr, c3psn = c_null_ptr, slatop = c_null_ptr, leafcn = c_null_ptr, flnr = c_null_ptr
     type(c_ptr) :: fnitr = c_null_ptr, mbbopt = c_null_ptr, medlynslope = c_null_ptr, medlynintercept = c_null_ptr, smpso = c_null_ptr
     type(c_ptr) :: smpsc = c_null_ptr, theta_cj = c_null_ptr, z0v_Cr = c_null_ptr, z0v_Cs = c_null_ptr, z0v_c = c_null_ptr
     type(c_ptr) :: z0v_cw = c_null_ptr, z0v_LAImax = c_null_ptr, is_tree = c_null_ptr, is_shrub = c_null_ptr, dbh = c_null_ptr
     type(c_ptr) :: fbw = c_null_ptr, nstem = c_null_ptr, rstem_per_dbh = c_null_ptr, wood_density = c_null_ptr
  end type ctsm_pft_params_t

  ! CanopyFluxes state (member order and meaning as ctsm_canopy_state)
  type, bind(C), public :: ctsm_canopy_state_t
     type(c_ptr) :: itype = c_null_ptr, column = c_null_ptr, gridcell = c_null_ptr, wtcol = c_null_ptr
     type(c_ptr) :: frac_veg_nosno = c_null_ptr, active = c_null_ptr, forc_u = c_null_ptr, forc_v = c_null_ptr
     type(c_ptr) :: forc_pco2 = c_null_ptr, forc_po2 = c_null_ptr, forc_hgt_u = c_null_ptr, forc_hgt_t = c_null_ptr
     type(c_ptr) :: forc_hgt_q = c_null_ptr, dayl = c_null_ptr, max_dayl = c_null_ptr, near_local_noon = c_null_ptr
     type(c_ptr) :: forc_q = c_null_ptr, forc_pbot = c_null_ptr, forc_th = c_null_ptr, forc_rho = c_null_ptr
     type(c_ptr) :: thv = c_null_ptr, qg = c_null_ptr, qg_snow = c_null_ptr, qg_soil = c_null_ptr
     type(c_ptr) :: qg_h2osfc = c_null_ptr, dqgdT = c_null_ptr, z0mg = c_null_ptr, soilbeta = c_null_ptr
     type(c_ptr) :: soilresis = c_null_ptr, h2osoi_liqvol = c_null_ptr, qflx_tran_veg_col = c_null_ptr, rootr_col = c_null_ptr
     type(c_ptr) :: qflx_rootsoi_col = c_null_ptr, elai = c_null_ptr, esai = c_null_ptr, laisun = c_null_ptr
     type(c_ptr) :: laisha = c_null_ptr, htop = c_null_ptr, sabv = c_null_ptr, parsun_z = c_null_ptr
     type(c_ptr) :: parsha_z = c_null_ptr, laisun_z = c_null_ptr, laisha_z = c_null_ptr, vcmaxcintsun = c_null_ptr
     type(c_ptr) :: vcmaxcintsha = c_null_ptr, nrad = c_null_ptr, thm = c_null_ptr, emv = c_null_ptr
     type(c_ptr) :: fwet = c_null_ptr, fdry = c_null_ptr, t_a10 = c_null_ptr, o3coefvsun = c_null_ptr
     type(c_ptr) :: o3coefvsha = c_null_ptr, o3coefgsun = c_null_ptr, o3coefgsha = c_null_ptr, rootfr = c_null_ptr
     type(c_ptr) :: displa = c_null_ptr, z0mv = c_null_ptr, t_veg = c_null_ptr, t_stem = c_null_ptr
     type(c_ptr) :: snocan = c_null_ptr, liqcan = c_null_ptr, cgrnds = c_null_ptr, cgrndl = c_null_ptr
     type(c_ptr) :: z0hv = c_null_ptr, z0qv = c_null_ptr, forc_hgt_u_patch = c_null_ptr, forc_hgt_t_patch = c_null_ptr
     type(c_ptr) :: forc_hgt_q_patch = c_null_ptr, ram1 = c_null_ptr, rb1 = c_null_ptr, rah1 = c_null_ptr
     type(c_ptr) :: rah2 = c_null_ptr, raw1 = c_null_ptr, raw2 = c_null_ptr, ustar = c_null_ptr
     type(c_ptr) :: um = c_null_ptr, uaf = c_null_ptr, taf = c_null_ptr, qaf = c_null_ptr
     type(c_ptr) :: obu = c_null_ptr, zeta = c_null_ptr, vpd = c_null_ptr, num_iter = c_null_ptr
     type(c_ptr) :: u10 = c_null_ptr, u10_clm = c_null_ptr, va = c_null_ptr, fv = c_null_ptr
     type(c_ptr) :: vds = c_null_ptr, btran = c_null_ptr, rresis = c_null_ptr, rootr = c_null_ptr
     type(c_ptr) :: eflx_sh_veg = c_null_ptr, eflx_sh_stem = c_null_ptr, eflx_sh_grnd = c_null_ptr, eflx_sh_snow = c_null_ptr
     type(c_ptr) :: eflx_sh_soil = c_null_ptr, eflx_sh_h2osfc = c_null_ptr, taux = c_null_ptr, tauy = c_null_ptr
     type(c_ptr) :: dlrad = c_null_ptr, ulrad = c_null_ptr, cgrnd = c_null_ptr, dhsdt_canopy = c_null_ptr
     type(c_ptr) :: qflx_evap_soi = c_null_ptr, qflx_ev_snow = c_null_ptr, qflx_ev_soil = c_null_ptr, qflx_ev_h2osfc = c_null_ptr
     type(c_ptr) :: qflx_tran_veg = c_null_ptr, qflx_evap_veg = c_null_ptr, rh_af = c_null_ptr, q_ref2m = c_null_ptr
     type(c_ptr) :: rh_ref2m = c_null_ptr, rh_ref2m_r = c_null_ptr, vpd_ref2m = c_null_ptr, iwue_ln = c_null_ptr
     type(c_ptr) :: t_ref2m = c_null_ptr, t_ref2m_r = c_null_ptr, t_skin = c_null_ptr, stem_biomass = c_null_ptr
     type(c_ptr) :: leaf_biomass = c_null_ptr, dleaf_patch = c_null_ptr, rssun = c_null_ptr, rssha = c_null_ptr
     type(c_ptr) :: psnsun = c_null_ptr, psnsha = c_null_ptr, psnsun_wc = c_null_ptr, psnsun_wj = c_null_ptr
     type(c_ptr) :: psnsun_wp = c_null_ptr, psnsha_wc = c_null_ptr, psnsha_wj = c_null_ptr, psnsha_wp = c_null_ptr
     type(c_ptr) :: fpsn = c_null_ptr, fpsn_wc = c_null_ptr, fpsn_wj = c_null_ptr, fpsn_wp = c_null_ptr
     type(c_ptr) :: lmrsun = c_null_ptr, lmrsha = c_null_ptr, gs_mol_sun = c_null_ptr, gs_mol_sha = c_null_ptr
     type(c_ptr) :: gs_mol_sun_ln = c_null_ptr, gs_mol_sha_ln = c_null_ptr, cisun_z = c_null_ptr, cisha_z = c_null_ptr
     type(c_ptr) :: rssun_z = c_null_ptr, rssha_z = c_null_ptr, lmrsun_z = c_null_ptr, lmrsha_z = c_null_ptr
     type(c_ptr) :: psnsun_z = c_null_ptr, psnsha_z = c_null_ptr, vcmax_z = c_null_ptr, tpu_z = c_null_ptr
     type(c_ptr) :: kp_z = c_null_ptr, ac = c_null_ptr, aj = c_null_ptr, ap = c_null_ptr
     type(c_ptr) :: ag = c_null_ptr, an = c_null_ptr, gb_mol = c_null_ptr, gs_mol = c_null_ptr
     type(c_ptr) :: cp = c_null_ptr, kc = c_null_ptr, ko = c_null_ptr, qe = c_null_ptr
     type(c_ptr) :: bbb = c_null_ptr, mbb = c_null_ptr, rh_leaf = c_null_ptr, vpd_can = c_null_ptr
     type(c_ptr) :: lnca = c_null_ptr, c3flag = c_null_ptr, canopy_err = c_null_ptr
  end type ctsm_canopy_state_t

  ! SoilFluxes state (member order and meaning as ctsm_soilflux_state)
  type, bind(C), public :: ctsm_soilflux_state_t
     type(c_ptr) :: active = c_null_ptr, t_ssbef = c_null_ptr, t_h2osfc_bef = c_null_ptr
     type(c_ptr) :: eflx_sh_grnd = c_null_ptr, qflx_evap_soi = c_null_ptr, qflx_ev_snow = c_null_ptr
     type(c_ptr) :: qflx_ev_soil = c_null_ptr, qflx_ev_h2osfc = c_null_ptr, cgrnds = c_null_ptr
     type(c_ptr) :: cgrndl = c_null_ptr, ulrad = c_null_ptr, eflx_sh_veg = c_null_ptr
     type(c_ptr) :: eflx_sh_stem = c_null_ptr, qflx_evap_veg = c_null_ptr, qflx_tran_veg = c_null_ptr
     type(c_ptr) :: eflx_soil_grnd = c_null_ptr, eflx_soil_grnd_r = c_null_ptr, eflx_sh_tot = c_null_ptr
     type(c_ptr) :: eflx_sh_tot_r = c_null_ptr, qflx_evap_tot = c_null_ptr, eflx_lh_tot = c_null_ptr
     type(c_ptr) :: eflx_lh_tot_r = c_null_ptr, qflx_evap_can = c_null_ptr, eflx_lh_vege = c_null_ptr
     type(c_ptr) :: eflx_lh_vegt = c_null_ptr, eflx_lh_grnd = c_null_ptr, eflx_lwrad_out = c_null_ptr
     type(c_ptr) :: eflx_lwrad_out_r = c_null_ptr, eflx_lwrad_net = c_null_ptr, eflx_lwrad_net_r = c_null_ptr
     type(c_ptr) :: t_skin = c_null_ptr, qflx_liqevap_from_top_layer = c_null_ptr, qflx_solidevap_from_top_layer = c_null_ptr
     type(c_ptr) :: qflx_liqdew_to_top_layer = c_null_ptr, qflx_soliddew_to_top_layer = c_null_ptr, errsoi_patch = c_null_ptr
     type(c_ptr) :: errsoi_col = c_null_ptr
  end type ctsm_soilflux_state_t

  ! Surface water input chain state (member order and meaning as ctsm_surface_water_state)
  type, bind(C), public :: ctsm_surface_water_state_t
     type(c_ptr) :: qflx_rain_plus_snomelt = c_null_ptr, qflx_snow_h2osfc = c_null_ptr, qflx_floodc = c_null_ptr
     type(c_ptr) :: qflx_liqevap_from_top_layer = c_null_ptr, qflx_ev_soil = c_null_ptr, qflx_ev_h2osfc = c_null_ptr
     type(c_ptr) :: zwt = c_null_ptr, zwt_perched = c_null_ptr, frost_table = c_null_ptr, wtfact = c_null_ptr
     type(c_ptr) :: topo_slope = c_null_ptr, h2osfc_thresh = c_null_ptr, frac_h2osfc_nosnow = c_null_ptr
     type(c_ptr) :: h2osfc = c_null_ptr
     type(c_ptr) :: fsat = c_null_ptr, fcov = c_null_ptr, qflx_sat_excess_surf = c_null_ptr
     type(c_ptr) :: qflx_top_soil = c_null_ptr, qflx_in_soil = c_null_ptr, qflx_top_soil_to_h2osfc = c_null_ptr
     type(c_ptr) :: qinmax = c_null_ptr, qflx_infl_excess = c_null_ptr, qflx_in_soil_limited = c_null_ptr
     type(c_ptr) :: qflx_in_h2osfc = c_null_ptr, qflx_infl_excess_surf = c_null_ptr
     type(c_ptr) :: qflx_h2osfc_surf = c_null_ptr, qflx_h2osfc_drain = c_null_ptr, qflx_infl = c_null_ptr
     type(c_ptr) :: qflx_surf = c_null_ptr
  end type ctsm_surface_water_state_t

  ! Water tables and condensation state (member order and meaning as ctsm_water_table_state)
  type, bind(C), public :: ctsm_water_table_state_t
     type(c_ptr) :: zwt = c_null_ptr, zwt_perched = c_null_ptr, frost_table = c_null_ptr
     type(c_ptr) :: h2osoi_vol = c_null_ptr
     type(c_ptr) :: qflx_liqdew_to_top_layer = c_null_ptr, qflx_soliddew_to_top_layer = c_null_ptr
     type(c_ptr) :: qflx_solidevap_from_top_layer = c_null_ptr
  end type ctsm_water_table_state_t

  ! Drainage state (member order and meaning as ctsm_drainage_state)
  type, bind(C), public :: ctsm_drainage_state_t
     type(c_ptr) :: topo_slope = c_null_ptr, zwt_perched = c_null_ptr, frost_table = c_null_ptr
     type(c_ptr) :: qflx_surf = c_null_ptr, qflx_snwcp_liq = c_null_ptr, zwt = c_null_ptr
     type(c_ptr) :: h2osoi_vol = c_null_ptr
     type(c_ptr) :: qflx_drain_perched = c_null_ptr, qflx_latflow_out = c_null_ptr, qflx_latflow_in = c_null_ptr
     type(c_ptr) :: qflx_rsub_sat = c_null_ptr, qflx_ice_runoff_xs = c_null_ptr, qflx_drain = c_null_ptr
     type(c_ptr) :: qflx_qrgwl = c_null_ptr, qflx_runoff = c_null_ptr
  end type ctsm_drainage_state_t

  ! column water stores and fluxes of the water balance check (member order and
  ! meaning as ctsm_water_balance_state); device arrays
  type, bind(C), public :: ctsm_water_balance_state_t
     type(c_ptr) :: total_plant_stored_h2o = c_null_ptr, wa = c_null_ptr
     type(c_ptr) :: excess_ice = c_null_ptr, dynbal_baseline_liq = c_null_ptr, dynbal_baseline_ice = c_null_ptr
     type(c_ptr) :: hydrologically_active = c_null_ptr, col_active = c_null_ptr
     type(c_ptr) :: forc_rain = c_null_ptr, forc_snow = c_null_ptr, qflx_floodc = c_null_ptr
     type(c_ptr) :: qflx_sfc_irrig = c_null_ptr, qflx_glcice_dyn_water_flux = c_null_ptr
     type(c_ptr) :: qflx_evap_tot = c_null_ptr, qflx_surf = c_null_ptr, qflx_qrgwl = c_null_ptr
     type(c_ptr) :: qflx_drain = c_null_ptr, qflx_drain_perched = c_null_ptr, qflx_ice_runoff = c_null_ptr
     type(c_ptr) :: qflx_snwcp_discarded_liq = c_null_ptr, qflx_snwcp_discarded_ice = c_null_ptr
     type(c_ptr) :: begwb = c_null_ptr, endwb = c_null_ptr, errh2o = c_null_ptr
  end type ctsm_water_balance_state_t

  ! fff, pc, mu, the drainage parameters (param file) and h2osfcflag, as ctsm_hydro_params
  type, bind(C), public :: ctsm_hydro_params_t
     real(c_double) :: fff, pc, mu
     real(c_double) :: aq_sp_yield_min, n_baseflow, baseflow_scalar, perched_baseflow_scalar
     integer(c_int32_t) :: h2osfcflag
     integer(c_int32_t) :: pad = 0
  end type ctsm_hydro_params_t

  ! plant hydraulic stress, soil side, as ctsm_phs_state: patch members, then
  ! column members; vegwp is vegwp_patch(begp:endp, 1:4)
  type, bind(C), public :: ctsm_phs_state_t
     type(c_ptr) :: itype = c_null_ptr, column = c_null_ptr, active = c_null_ptr, frac_veg_nosno = c_null_ptr
     type(c_ptr) :: wtcol = c_null_ptr, froot_carbon = c_null_ptr, tsai = c_null_ptr, tlai = c_null_ptr
     type(c_ptr) :: rootfr = c_null_ptr, vegwp = c_null_ptr, k_soil_root = c_null_ptr
     type(c_ptr) :: root_conductance = c_null_ptr, soil_conductance = c_null_ptr, qflx_hydr_redist = c_null_ptr
     type(c_ptr) :: patchi = c_null_ptr, patchf = c_null_ptr, dz = c_null_ptr, z = c_null_ptr
     type(c_ptr) :: hksat = c_null_ptr, hk_l = c_null_ptr, smp_l = c_null_ptr, qflx_rootsoi = c_null_ptr
     type(c_ptr) :: qflx_phs_neg = c_null_ptr
  end type ctsm_phs_state_t

  ! PHS tables (params_inst%krmax, psi50, ck; pftcon%root_radius, root_density,
  ! froot_leaf), psi50 and ck as (0:npft-1, 1:4) arrays
  type, bind(C), public :: ctsm_phs_params_t
     integer(c_int32_t) :: npft
     integer(c_int32_t) :: pad = 0
     type(c_ptr) :: krmax = c_null_ptr, root_radius = c_null_ptr, root_density = c_null_ptr
     type(c_ptr) :: froot_leaf = c_null_ptr, psi50 = c_null_ptr, ck = c_null_ptr
     type(c_ptr) :: kmax = c_null_ptr, crop = c_null_ptr
  end type ctsm_phs_params_t

  ! CanopyFluxes with plant hydraulic stress, as ctsm_phs_canopy_state
  ! ([VEG] (begp:endp, 1:4), [PHS2] (begp:endp, 1:2))
  type, bind(C), public :: ctsm_phs_canopy_state_t
     type(c_ptr) :: tlai_z = c_null_ptr, local_morning = c_null_ptr
     type(c_ptr) :: vegwp = c_null_ptr, vegwp_ln = c_null_ptr, vegwp_pd = c_null_ptr
     type(c_ptr) :: bsun = c_null_ptr, bsha = c_null_ptr, an_sun = c_null_ptr, an_sha = c_null_ptr
     type(c_ptr) :: vcmax_z_phs = c_null_ptr, tpu_z_phs = c_null_ptr, kp_z_phs = c_null_ptr
     type(c_ptr) :: ac_phs = c_null_ptr, aj_phs = c_null_ptr, ap_phs = c_null_ptr, ag_phs = c_null_ptr
  end type ctsm_phs_canopy_state_t

  ! calcstress arrays, as ctsm_phs_stress_state (vegwp / vegwp_pd (begp:endp, 1:4))
  type, bind(C), public :: ctsm_phs_stress_state_t
     type(c_ptr) :: itype = c_null_ptr, column = c_null_ptr, local_morning = c_null_ptr
     type(c_ptr) :: laisun = c_null_ptr, laisha = c_null_ptr, elai = c_null_ptr, esai = c_null_ptr
     type(c_ptr) :: tsai = c_null_ptr, htop = c_null_ptr, fdry = c_null_ptr, thm = c_null_ptr
     type(c_ptr) :: gb_mol = c_null_ptr, gs_mol_sun = c_null_ptr, gs_mol_sha = c_null_ptr
     type(c_ptr) :: qsatl = c_null_ptr, qaf = c_null_ptr, k_soil_root = c_null_ptr
     type(c_ptr) :: vegwp = c_null_ptr, vegwp_pd = c_null_ptr, bsun = c_null_ptr, bsha = c_null_ptr
     type(c_ptr) :: qflx_tran_veg = c_null_ptr, forc_rho = c_null_ptr, forc_pbot = c_null_ptr
     type(c_ptr) :: z = c_null_ptr, smp_l = c_null_ptr
  end type ctsm_phs_stress_state_t

  public :: ctsm_config_default, ctsm_create, ctsm_destroy, ctsm_last_error
  public :: ctsm_soilwater, ctsm_soiltemperature, ctsm_soilwater_fractions
  public :: ctsm_set_canopy_params, ctsm_canopyfluxes, ctsm_root_sink
  public :: ctsm_soilfluxes, ctsm_soil_balance_check, ctsm_surface_water, ctsm_water_table, ctsm_drainage
  public :: ctsm_set_phs_params, ctsm_phs_soil_root_conductance, ctsm_phs_root_sink, ctsm_phs_calcstress

  interface
     subroutine ctsm_config_default(cfg) bind(C, name='ctsm_config_default')
       import :: ctsm_config_t
       type(ctsm_config_t), intent(out) :: cfg
     end subroutine ctsm_config_default

     integer(c_int) function ctsm_create(cfg, device, ctx) bind(C, name='ctsm_create')
       import :: c_int, c_ptr, ctsm_config_t
       type(ctsm_config_t), intent(in) :: cfg
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: ctx
     end function ctsm_create

     integer(c_int) function ctsm_destroy(ctx) bind(C, name='ctsm_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function ctsm_destroy

     integer(c_int) function ctsm_last_error(ctx, code, index) bind(C, name='ctsm_last_error')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: code, index
     end function ctsm_last_error

     integer(c_int) function ctsm_soilwater(ctx, lay, col, num_hydrologyc, filter_hydrologyc, dtime) &
          bind(C, name='ctsm_soilwater')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       integer(c_int32_t), value :: num_hydrologyc
       integer(c_int32_t), intent(in) :: filter_hydrologyc(*)
       real(c_double), value :: dtime
     end function ctsm_soilwater

     integer(c_int) function ctsm_soiltemperature(ctx, lay, col, pch, num_nolakec, filter_nolakec, dtime) &
          bind(C, name='ctsm_soiltemperature')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_patch_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_patch_state_t), intent(in) :: pch
       integer(c_int32_t), value :: num_nolakec
       integer(c_int32_t), intent(in) :: filter_nolakec(*)
       real(c_double), value :: dtime
     end function ctsm_soiltemperature

     integer(c_int) function ctsm_soilwater_fractions(ctx, lay, col, num_hydrologyc, filter_hydrologyc) &
          bind(C, name='ctsm_soilwater_fractions')
       import :: c_int, c_int32_t, c_ptr, ctsm_layout_t, ctsm_col_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       integer(c_int32_t), value :: num_hydrologyc
       integer(c_int32_t), intent(in) :: filter_hydrologyc(*)
     end function ctsm_soilwater_fractions

     integer(c_int) function ctsm_set_canopy_params(ctx, prm, pft) bind(C, name='ctsm_set_canopy_params')
       import :: c_int, c_ptr, ctsm_canopy_params_t, ctsm_pft_params_t
       type(c_ptr), value :: ctx
       type(ctsm_canopy_params_t), intent(in) :: prm
       type(ctsm_pft_params_t), intent(in) :: pft
     end function ctsm_set_canopy_params

     integer(c_int) function ctsm_canopyfluxes(ctx, lay, col, can, num_exposedvegp, filter_exposedvegp, dtime) &
          bind(C, name='ctsm_canopyfluxes')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_canopy_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_canopy_state_t), intent(in) :: can
       integer(c_int32_t), value :: num_exposedvegp
       integer(c_int32_t), intent(in) :: filter_exposedvegp(*)
       real(c_double), value :: dtime
     end function ctsm_canopyfluxes

     integer(c_int) function ctsm_root_sink(ctx, lay, col, can, num_filterc, filterc) bind(C, name='ctsm_root_sink')
       import :: c_int, c_int32_t, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_canopy_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_canopy_state_t), intent(in) :: can
       integer(c_int32_t), value :: num_filterc
       integer(c_int32_t), intent(in) :: filterc(*)
     end function ctsm_root_sink

     integer(c_int) function ctsm_soilfluxes(ctx, lay, col, pch, sf, num_nolakec, filter_nolakec, &
          num_nolakep, filter_nolakep, dtime) bind(C, name='ctsm_soilfluxes')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_patch_state_t, &
            ctsm_soilflux_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_patch_state_t), intent(in) :: pch
       type(ctsm_soilflux_state_t), intent(in) :: sf
       integer(c_int32_t), value :: num_nolakec
       integer(c_int32_t), intent(in) :: filter_nolakec(*)
       integer(c_int32_t), value :: num_nolakep
       integer(c_int32_t), intent(in) :: filter_nolakep(*)
       real(c_double), value :: dtime
     end function ctsm_soilfluxes

     integer(c_int) function ctsm_soil_balance_check(ctx, lay, sf, num_filterc, filterc, errsoi_max, idx) &
          bind(C, name='ctsm_soil_balance_check')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_soilflux_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_soilflux_state_t), intent(in) :: sf
       integer(c_int32_t), value :: num_filterc
       integer(c_int32_t), intent(in) :: filterc(*)
       real(c_double), intent(out) :: errsoi_max
       integer(c_int32_t), intent(out) :: idx
     end function ctsm_soil_balance_check

     integer(c_int) function ctsm_surface_water(ctx, lay, col, sw, prm, num_hydrologyc, filter_hydrologyc, &
          dtime) bind(C, name='ctsm_surface_water')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, &
            ctsm_surface_water_state_t, ctsm_hydro_params_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_surface_water_state_t), intent(in) :: sw
       type(ctsm_hydro_params_t), intent(in) :: prm
       integer(c_int32_t), value :: num_hydrologyc
       integer(c_int32_t), intent(in) :: filter_hydrologyc(*)
       real(c_double), value :: dtime
     end function ctsm_surface_water

     integer(c_int) function ctsm_water_table(ctx, lay, col, wt, num_hydrologyc, filter_hydrologyc, dtime) &
          bind(C, name='ctsm_water_table')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_water_table_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_water_table_state_t), intent(in) :: wt
       integer(c_int32_t), value :: num_hydrologyc
       integer(c_int32_t), intent(in) :: filter_hydrologyc(*)
       real(c_double), value :: dtime
     end function ctsm_water_table

     integer(c_int) function ctsm_drainage(ctx, lay, col, dr, prm, num_hydrologyc, filter_hydrologyc, dtime) &
          bind(C, name='ctsm_drainage')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_drainage_state_t, &
            ctsm_hydro_params_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_drainage_state_t), intent(in) :: dr
       type(ctsm_hydro_params_t), intent(in) :: prm
       integer(c_int32_t), value :: num_hydrologyc
       integer(c_int32_t), intent(in) :: filter_hydrologyc(*)
       real(c_double), value :: dtime
     end function ctsm_drainage

     integer(c_int) function ctsm_set_phs_params(ctx, prm) bind(C, name='ctsm_set_phs_params')
       import :: c_int, c_ptr, ctsm_phs_params_t
       type(c_ptr), value :: ctx
       type(ctsm_phs_params_t), intent(in) :: prm
     end function ctsm_set_phs_params

     integer(c_int) function ctsm_phs_soil_root_conductance(ctx, lay, phs, num_filterp, filterp) &
          bind(C, name='ctsm_phs_soil_root_conductance')
       import :: c_int, c_int32_t, c_ptr, ctsm_layout_t, ctsm_phs_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_phs_state_t), intent(in) :: phs
       integer(c_int32_t), value :: num_filterp
       integer(c_int32_t), intent(in) :: filterp(*)
     end function ctsm_phs_soil_root_conductance

     integer(c_int) function ctsm_phs_root_sink(ctx, lay, phs, num_filterc, filterc) &
          bind(C, name='ctsm_phs_root_sink')
       import :: c_int, c_int32_t, c_ptr, ctsm_layout_t, ctsm_phs_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_phs_state_t), intent(in) :: phs
       integer(c_int32_t), value :: num_filterc
       integer(c_int32_t), intent(in) :: filterc(*)
     end function ctsm_phs_root_sink

     integer(c_int) function ctsm_phs_calcstress(ctx, lay, st, num_filterp, filterp) &
          bind(C, name='ctsm_phs_calcstress')
       import :: c_int, c_int32_t, c_ptr, ctsm_layout_t, ctsm_phs_stress_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_phs_stress_state_t), intent(in) :: st
       integer(c_int32_t), value :: num_filterp
       integer(c_int32_t), intent(in) :: filterp(*)
     end function ctsm_phs_calcstress

     ! CanopyFluxes with use_hydrstress = .true. (the context's config has
     ! use_hydrstress = 1): in place of the call at clm_driver.F90:766
     integer(c_int) function ctsm_canopyfluxes_phs(ctx, lay, col, can, phs, phc, num_exposedvegp, &
          filter_exposedvegp, dtime) bind(C, name='ctsm_canopyfluxes_phs')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_canopy_state_t, &
            ctsm_phs_state_t, ctsm_phs_canopy_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_canopy_state_t), intent(in) :: can
       type(ctsm_phs_state_t), intent(in) :: phs
       type(ctsm_phs_canopy_state_t), intent(in) :: phc
       integer(c_int32_t), value :: num_exposedvegp
       integer(c_int32_t), intent(in) :: filter_exposedvegp(*)
       real(c_double), value :: dtime
     end function ctsm_canopyfluxes_phs

     ! ComputeWaterMassNonLake for begwb / endwb on device-resident state
     ! (TotalWaterAndHeatMod.F90:92-393); water_mass is a device (c) array
     integer(c_int) function ctsm_water_mass_dev(ctx, lay, col, can, wb, num_nolakec, filter_nolakec_dev, &
          aquifer_water_baseline, subtract_dynbal_baselines, water_mass) bind(C, name='ctsm_water_mass_dev')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_col_state_t, ctsm_canopy_state_t, &
            ctsm_water_balance_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_col_state_t), intent(in) :: col
       type(ctsm_canopy_state_t), intent(in) :: can
       type(ctsm_water_balance_state_t), intent(in) :: wb
       integer(c_int32_t), value :: num_nolakec
       type(c_ptr), value :: filter_nolakec_dev
       real(c_double), value :: aquifer_water_baseline
       integer(c_int32_t), value :: subtract_dynbal_baselines
       type(c_ptr), value :: water_mass
     end function ctsm_water_mass_dev

     ! the column water balance check of BalanceCheck (BalanceCheckMod.F90:588-630)
     integer(c_int) function ctsm_water_balance_check_dev(ctx, lay, wb, num_filterc, filterc_dev, dtime, &
          errh2o_max, errh2o_col) bind(C, name='ctsm_water_balance_check_dev')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_water_balance_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_water_balance_state_t), intent(in) :: wb
       integer(c_int32_t), value :: num_filterc
       type(c_ptr), value :: filterc_dev
       real(c_double), value :: dtime
       real(c_double), intent(out) :: errh2o_max
       integer(c_int32_t), intent(out) :: errh2o_col
     end function ctsm_water_balance_check_dev

     ! the solar radiation balance of EnergyBalanceCheck (BalanceCheckMod.F90:964-970, 1008-1012)
     integer(c_int) function ctsm_solar_balance_check_dev(ctx, lay, can, num_filterp, filterp_dev, fsa, fsr, &
          forc_solad_col, forc_solai, errsol, errsol_max, sol_patch) bind(C, name='ctsm_solar_balance_check_dev')
       import :: c_int, c_int32_t, c_double, c_ptr, ctsm_layout_t, ctsm_canopy_state_t
       type(c_ptr), value :: ctx
       type(ctsm_layout_t), intent(in) :: lay
       type(ctsm_canopy_state_t), intent(in) :: can
       integer(c_int32_t), value :: num_filterp
       type(c_ptr), value :: filterp_dev, fsa, fsr, forc_solad_col, forc_solai, errsol
       real(c_double), intent(out) :: errsol_max
       integer(c_int32_t), intent(out) :: sol_patch
     end function ctsm_solar_balance_check_dev
  end interface

end module ctsm_gpu_mod
