module ctsm_gpu_mod
  !-----------------------------------------------------------------------
  ! ISO_C_BINDING interface of libctsm_mi355x.so (include/ctsm_gpu.h).
  !
  ! The derived types below are interoperable mirrors of the C structs; the
  ! state structs hold c_ptr members that the caller sets with c_loc() of
  ! the reference's own arrays (e.g. c_loc(waterstatebulk_inst%h2osoi_liq_col)),
  ! so the library reads and writes CTSM's memory in place.
  !
  ! Drop-in wrappers for SoilWater (SoilWaterMovementMod.F90:240-243) and
  ! SoilTemperature (SoilTemperatureMod.F90:92-95) built on this module are
  ! given in INTEGRATION.md; they fill the structs from the *_inst objects,
  ! convert nothing (filters stay 1-based reference indices) and call endrun
  ! with the failing column when a call returns CTSM_E_SOLVER.
  !-----------------------------------------------------------------------
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CTSM_OK = 0, CTSM_E_ARG = 1, CTSM_E_CONFIG = 2, &
       CTSM_E_HIP = 3, CTSM_E_SOLVER = 4, CTSM_E_NODEVICE = 5
  integer(c_int), parameter, public :: CTSM_BC_FLUX = 1, CTSM_BC_ZERO_FLUX = 2
  integer(c_int), parameter, public :: CTSM_SNOWTHK_JORDAN1991 = 0, CTSM_SNOWTHK_STURM1997 = 1

  type, bind(C), public :: ctsm_config_t
     integer(c_int32_t) :: nlevsno, nlevsoi, nlevgrnd
     integer(c_int32_t) :: soilwater_movement_method
     integer(c_int32_t) :: upper_boundary_condition, lower_boundary_condition
     integer(c_int32_t) :: flux_calculation, snow_thermal_cond_method
     integer(c_int32_t) :: itmax_canopy_fluxes, pad0
     real(c_double)     :: dtmin, verySmall, xTolerUpper, xTolerLower, e_ice
  end type ctsm_config_t

  type, bind(C), public :: ctsm_layout_t
     integer(c_int32_t) :: begc_alloc, ldc, begp_alloc, ldp
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

  public :: ctsm_config_default, ctsm_create, ctsm_destroy, ctsm_last_error
  public :: ctsm_soilwater, ctsm_soiltemperature, ctsm_soilwater_fractions

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
  end interface

end module ctsm_gpu_mod
