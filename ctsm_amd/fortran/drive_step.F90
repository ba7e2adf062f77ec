! drive_step.F90 -- written by gen_drive_step.py from the member lists of
! ctsm_amd/abi.py; do not edit by hand.
!
! One of SoilWater, SoilTemperature or CanopyFluxes from Fortran through
! ctsm_gpu_mod, on a state held the way CTSM holds it: one allocatable array per
! member with the reference's bounds, bound by c_loc of its first element;
! canopy members that are column or patch arrays point at those arrays.
!
! Usage: drive_step IN OUT.  IN holds, as int32, ncol, npatch, ngrid, nlevsno,
! nlevsoi, nlevgrnd, mode (1 SoilWater, 2 SoilTemperature, 3 CanopyFluxes,
! 4 surface water, 5 water tables, 6 drainage) and npft; the ctsm_config bytes;
! the canopy parameters (doubles), the PFT table (npft values per field: the
! real fields, then the integer ones, each in member order) and the hydrology
! parameters; then every column, patch, canopy and hydrology array in member
! order.
! OUT receives the arrays after the call.  Prints the return code.
program drive_step
  use iso_c_binding
  use ctsm_gpu_mod
  implicit none
  integer(c_int32_t) :: ncol, npatch, ngrid, nlevsno, nlevsoi, nlevgrnd, mode, npft
  integer :: begc, endc, begp, endp, begg, endg, u, i, rc
  character(len=4096) :: fin, fout
  type(ctsm_config_t) :: cfg
  type(ctsm_layout_t) :: lay
  type(ctsm_col_state_t) :: col
  type(ctsm_patch_state_t) :: pch
  type(ctsm_canopy_state_t) :: can
  type(ctsm_canopy_params_t) :: prm
  type(ctsm_pft_params_t) :: pft
  type(ctsm_surface_water_state_t) :: sw
  type(ctsm_water_table_state_t) :: wt
  type(ctsm_drainage_state_t) :: dr
  type(ctsm_hydro_params_t) :: hprm
  type(c_ptr) :: ctx
  integer(c_int32_t), allocatable :: filter(:)
  real(c_double), allocatable, target :: pftbuf(:,:)
  integer(c_int32_t), allocatable, target :: pftibuf(:,:)
  integer(c_int32_t), allocatable, target :: c_snl(:)
  integer(c_int32_t), allocatable, target :: c_nbedrock(:)
  integer(c_int32_t), allocatable, target :: c_patchi(:)
  integer(c_int32_t), allocatable, target :: c_patchf(:)
  real(c_double), allocatable, target :: c_dz(:,:)
  real(c_double), allocatable, target :: c_z(:,:)
  real(c_double), allocatable, target :: c_zi(:,:)
  real(c_double), allocatable, target :: c_watsat(:,:)
  real(c_double), allocatable, target :: c_bsw(:,:)
  real(c_double), allocatable, target :: c_sucsat(:,:)
  real(c_double), allocatable, target :: c_hksat(:,:)
  real(c_double), allocatable, target :: c_tkmg(:,:)
  real(c_double), allocatable, target :: c_tkdry(:,:)
  real(c_double), allocatable, target :: c_csol(:,:)
  real(c_double), allocatable, target :: c_eff_porosity(:,:)
  real(c_double), allocatable, target :: c_smp_l(:,:)
  real(c_double), allocatable, target :: c_hk_l(:,:)
  real(c_double), allocatable, target :: c_thk(:,:)
  real(c_double), allocatable, target :: c_icefrac(:,:)
  real(c_double), allocatable, target :: c_qcharge(:)
  real(c_double), allocatable, target :: c_num_substeps(:)
  real(c_double), allocatable, target :: c_qin(:,:)
  real(c_double), allocatable, target :: c_qout(:,:)
  real(c_double), allocatable, target :: c_h2osoi_liq(:,:)
  real(c_double), allocatable, target :: c_h2osoi_ice(:,:)
  real(c_double), allocatable, target :: c_h2osfc(:)
  real(c_double), allocatable, target :: c_h2osno_no_layers(:)
  real(c_double), allocatable, target :: c_int_snow(:)
  real(c_double), allocatable, target :: c_frac_sno_eff(:)
  real(c_double), allocatable, target :: c_frac_h2osfc(:)
  real(c_double), allocatable, target :: c_snow_depth(:)
  real(c_double), allocatable, target :: c_bw(:,:)
  real(c_double), allocatable, target :: c_snomelt_accum(:)
  real(c_double), allocatable, target :: c_qflx_infl(:)
  real(c_double), allocatable, target :: c_qflx_rootsoi(:,:)
  real(c_double), allocatable, target :: c_qflx_snomelt(:)
  real(c_double), allocatable, target :: c_qflx_snofrz(:)
  real(c_double), allocatable, target :: c_qflx_snow_drain(:)
  real(c_double), allocatable, target :: c_qflx_snomelt_lyr(:,:)
  real(c_double), allocatable, target :: c_qflx_snofrz_lyr(:,:)
  real(c_double), allocatable, target :: c_qflx_h2osfc_to_ice(:)
  real(c_double), allocatable, target :: c_t_soisno(:,:)
  real(c_double), allocatable, target :: c_t_h2osfc(:)
  real(c_double), allocatable, target :: c_t_grnd(:)
  real(c_double), allocatable, target :: c_fact(:,:)
  real(c_double), allocatable, target :: c_xmf(:)
  real(c_double), allocatable, target :: c_xmf_h2osfc(:)
  real(c_double), allocatable, target :: c_c_h2osfc(:)
  integer(c_int32_t), allocatable, target :: c_imelt(:,:)
  real(c_double), allocatable, target :: c_emg(:)
  real(c_double), allocatable, target :: c_htvp(:)
  real(c_double), allocatable, target :: c_eflx_bot(:)
  real(c_double), allocatable, target :: c_eflx_fgr12(:)
  real(c_double), allocatable, target :: c_eflx_fgr(:,:)
  real(c_double), allocatable, target :: c_eflx_snomelt(:)
  real(c_double), allocatable, target :: c_eflx_h2osfc_to_snow(:)
  real(c_double), allocatable, target :: c_forc_lwrad(:)
  real(c_double), allocatable, target :: c_zwt(:)
  real(c_double), allocatable, target :: c_smpmin(:)
  integer(c_int32_t), allocatable, target :: p_column(:)
  real(c_double), allocatable, target :: p_wtcol(:)
  integer(c_int32_t), allocatable, target :: p_frac_veg_nosno(:)
  real(c_double), allocatable, target :: p_sabg(:)
  real(c_double), allocatable, target :: p_sabg_soil(:)
  real(c_double), allocatable, target :: p_sabg_snow(:)
  real(c_double), allocatable, target :: p_sabg_lyr(:,:)
  real(c_double), allocatable, target :: p_dlrad(:)
  real(c_double), allocatable, target :: p_cgrnd(:)
  real(c_double), allocatable, target :: p_eflx_sh_grnd(:)
  real(c_double), allocatable, target :: p_eflx_sh_snow(:)
  real(c_double), allocatable, target :: p_eflx_sh_soil(:)
  real(c_double), allocatable, target :: p_eflx_sh_h2osfc(:)
  real(c_double), allocatable, target :: p_qflx_evap_soi(:)
  real(c_double), allocatable, target :: p_qflx_ev_snow(:)
  real(c_double), allocatable, target :: p_qflx_ev_soil(:)
  real(c_double), allocatable, target :: p_qflx_ev_h2osfc(:)
  real(c_double), allocatable, target :: p_eflx_gnet(:)
  real(c_double), allocatable, target :: p_dgnetdT(:)
  real(c_double), allocatable, target :: p_sabg_chk(:)
  integer(c_int32_t), allocatable, target :: g_itype(:)
  integer(c_int32_t), allocatable, target :: g_gridcell(:)
  integer(c_int32_t), allocatable, target :: g_active(:)
  real(c_double), allocatable, target :: g_forc_u(:)
  real(c_double), allocatable, target :: g_forc_v(:)
  real(c_double), allocatable, target :: g_forc_pco2(:)
  real(c_double), allocatable, target :: g_forc_po2(:)
  real(c_double), allocatable, target :: g_forc_hgt_u(:)
  real(c_double), allocatable, target :: g_forc_hgt_t(:)
  real(c_double), allocatable, target :: g_forc_hgt_q(:)
  real(c_double), allocatable, target :: g_dayl(:)
  real(c_double), allocatable, target :: g_max_dayl(:)
  integer(c_int32_t), allocatable, target :: g_near_local_noon(:)
  real(c_double), allocatable, target :: g_forc_q(:)
  real(c_double), allocatable, target :: g_forc_pbot(:)
  real(c_double), allocatable, target :: g_forc_th(:)
  real(c_double), allocatable, target :: g_forc_rho(:)
  real(c_double), allocatable, target :: g_thv(:)
  real(c_double), allocatable, target :: g_qg(:)
  real(c_double), allocatable, target :: g_qg_snow(:)
  real(c_double), allocatable, target :: g_qg_soil(:)
  real(c_double), allocatable, target :: g_qg_h2osfc(:)
  real(c_double), allocatable, target :: g_dqgdT(:)
  real(c_double), allocatable, target :: g_z0mg(:)
  real(c_double), allocatable, target :: g_soilbeta(:)
  real(c_double), allocatable, target :: g_soilresis(:)
  real(c_double), allocatable, target :: g_h2osoi_liqvol(:,:)
  real(c_double), allocatable, target :: g_qflx_tran_veg_col(:)
  real(c_double), allocatable, target :: g_rootr_col(:,:)
  real(c_double), allocatable, target :: g_elai(:)
  real(c_double), allocatable, target :: g_esai(:)
  real(c_double), allocatable, target :: g_laisun(:)
  real(c_double), allocatable, target :: g_laisha(:)
  real(c_double), allocatable, target :: g_htop(:)
  real(c_double), allocatable, target :: g_sabv(:)
  real(c_double), allocatable, target :: g_parsun_z(:)
  real(c_double), allocatable, target :: g_parsha_z(:)
  real(c_double), allocatable, target :: g_laisun_z(:)
  real(c_double), allocatable, target :: g_laisha_z(:)
  real(c_double), allocatable, target :: g_vcmaxcintsun(:)
  real(c_double), allocatable, target :: g_vcmaxcintsha(:)
  integer(c_int32_t), allocatable, target :: g_nrad(:)
  real(c_double), allocatable, target :: g_thm(:)
  real(c_double), allocatable, target :: g_emv(:)
  real(c_double), allocatable, target :: g_fwet(:)
  real(c_double), allocatable, target :: g_fdry(:)
  real(c_double), allocatable, target :: g_t_a10(:)
  real(c_double), allocatable, target :: g_o3coefvsun(:)
  real(c_double), allocatable, target :: g_o3coefvsha(:)
  real(c_double), allocatable, target :: g_o3coefgsun(:)
  real(c_double), allocatable, target :: g_o3coefgsha(:)
  real(c_double), allocatable, target :: g_rootfr(:,:)
  real(c_double), allocatable, target :: g_displa(:)
  real(c_double), allocatable, target :: g_z0mv(:)
  real(c_double), allocatable, target :: g_t_veg(:)
  real(c_double), allocatable, target :: g_t_stem(:)
  real(c_double), allocatable, target :: g_snocan(:)
  real(c_double), allocatable, target :: g_liqcan(:)
  real(c_double), allocatable, target :: g_cgrnds(:)
  real(c_double), allocatable, target :: g_cgrndl(:)
  real(c_double), allocatable, target :: g_z0hv(:)
  real(c_double), allocatable, target :: g_z0qv(:)
  real(c_double), allocatable, target :: g_forc_hgt_u_patch(:)
  real(c_double), allocatable, target :: g_forc_hgt_t_patch(:)
  real(c_double), allocatable, target :: g_forc_hgt_q_patch(:)
  real(c_double), allocatable, target :: g_ram1(:)
  real(c_double), allocatable, target :: g_rb1(:)
  real(c_double), allocatable, target :: g_rah1(:)
  real(c_double), allocatable, target :: g_rah2(:)
  real(c_double), allocatable, target :: g_raw1(:)
  real(c_double), allocatable, target :: g_raw2(:)
  real(c_double), allocatable, target :: g_ustar(:)
  real(c_double), allocatable, target :: g_um(:)
  real(c_double), allocatable, target :: g_uaf(:)
  real(c_double), allocatable, target :: g_taf(:)
  real(c_double), allocatable, target :: g_qaf(:)
  real(c_double), allocatable, target :: g_obu(:)
  real(c_double), allocatable, target :: g_zeta(:)
  real(c_double), allocatable, target :: g_vpd(:)
  real(c_double), allocatable, target :: g_num_iter(:)
  real(c_double), allocatable, target :: g_u10(:)
  real(c_double), allocatable, target :: g_u10_clm(:)
  real(c_double), allocatable, target :: g_va(:)
  real(c_double), allocatable, target :: g_fv(:)
  real(c_double), allocatable, target :: g_vds(:)
  real(c_double), allocatable, target :: g_btran(:)
  real(c_double), allocatable, target :: g_rresis(:,:)
  real(c_double), allocatable, target :: g_rootr(:,:)
  real(c_double), allocatable, target :: g_eflx_sh_veg(:)
  real(c_double), allocatable, target :: g_eflx_sh_stem(:)
  real(c_double), allocatable, target :: g_taux(:)
  real(c_double), allocatable, target :: g_tauy(:)
  real(c_double), allocatable, target :: g_ulrad(:)
  real(c_double), allocatable, target :: g_dhsdt_canopy(:)
  real(c_double), allocatable, target :: g_qflx_tran_veg(:)
  real(c_double), allocatable, target :: g_qflx_evap_veg(:)
  real(c_double), allocatable, target :: g_rh_af(:)
  real(c_double), allocatable, target :: g_q_ref2m(:)
  real(c_double), allocatable, target :: g_rh_ref2m(:)
  real(c_double), allocatable, target :: g_rh_ref2m_r(:)
  real(c_double), allocatable, target :: g_vpd_ref2m(:)
  real(c_double), allocatable, target :: g_iwue_ln(:)
  real(c_double), allocatable, target :: g_t_ref2m(:)
  real(c_double), allocatable, target :: g_t_ref2m_r(:)
  real(c_double), allocatable, target :: g_t_skin(:)
  real(c_double), allocatable, target :: g_stem_biomass(:)
  real(c_double), allocatable, target :: g_leaf_biomass(:)
  real(c_double), allocatable, target :: g_dleaf_patch(:)
  real(c_double), allocatable, target :: g_rssun(:)
  real(c_double), allocatable, target :: g_rssha(:)
  real(c_double), allocatable, target :: g_psnsun(:)
  real(c_double), allocatable, target :: g_psnsha(:)
  real(c_double), allocatable, target :: g_psnsun_wc(:)
  real(c_double), allocatable, target :: g_psnsun_wj(:)
  real(c_double), allocatable, target :: g_psnsun_wp(:)
  real(c_double), allocatable, target :: g_psnsha_wc(:)
  real(c_double), allocatable, target :: g_psnsha_wj(:)
  real(c_double), allocatable, target :: g_psnsha_wp(:)
  real(c_double), allocatable, target :: g_fpsn(:)
  real(c_double), allocatable, target :: g_fpsn_wc(:)
  real(c_double), allocatable, target :: g_fpsn_wj(:)
  real(c_double), allocatable, target :: g_fpsn_wp(:)
  real(c_double), allocatable, target :: g_lmrsun(:)
  real(c_double), allocatable, target :: g_lmrsha(:)
  real(c_double), allocatable, target :: g_gs_mol_sun(:)
  real(c_double), allocatable, target :: g_gs_mol_sha(:)
  real(c_double), allocatable, target :: g_gs_mol_sun_ln(:)
  real(c_double), allocatable, target :: g_gs_mol_sha_ln(:)
  real(c_double), allocatable, target :: g_cisun_z(:)
  real(c_double), allocatable, target :: g_cisha_z(:)
  real(c_double), allocatable, target :: g_rssun_z(:)
  real(c_double), allocatable, target :: g_rssha_z(:)
  real(c_double), allocatable, target :: g_lmrsun_z(:)
  real(c_double), allocatable, target :: g_lmrsha_z(:)
  real(c_double), allocatable, target :: g_psnsun_z(:)
  real(c_double), allocatable, target :: g_psnsha_z(:)
  real(c_double), allocatable, target :: g_vcmax_z(:)
  real(c_double), allocatable, target :: g_tpu_z(:)
  real(c_double), allocatable, target :: g_kp_z(:)
  real(c_double), allocatable, target :: g_ac(:)
  real(c_double), allocatable, target :: g_aj(:)
  real(c_double), allocatable, target :: g_ap(:)
  real(c_double), allocatable, target :: g_ag(:)
  real(c_double), allocatable, target :: g_an(:)
  real(c_double), allocatable, target :: g_gb_mol(:)
  real(c_double), allocatable, target :: g_gs_mol(:)
  real(c_double), allocatable, target :: g_cp(:)
  real(c_double), allocatable, target :: g_kc(:)
  real(c_double), allocatable, target :: g_ko(:)
  real(c_double), allocatable, target :: g_qe(:)
  real(c_double), allocatable, target :: g_bbb(:)
  real(c_double), allocatable, target :: g_mbb(:)
  real(c_double), allocatable, target :: g_rh_leaf(:)
  real(c_double), allocatable, target :: g_vpd_can(:)
  real(c_double), allocatable, target :: g_lnca(:)
  integer(c_int32_t), allocatable, target :: g_c3flag(:)
  real(c_double), allocatable, target :: g_canopy_err(:)
  real(c_double), allocatable, target :: y_qflx_rain_plus_snomelt(:)
  real(c_double), allocatable, target :: y_qflx_snow_h2osfc(:)
  real(c_double), allocatable, target :: y_qflx_floodc(:)
  real(c_double), allocatable, target :: y_qflx_liqevap_from_top_layer(:)
  real(c_double), allocatable, target :: y_qflx_ev_soil(:)
  real(c_double), allocatable, target :: y_qflx_ev_h2osfc(:)
  real(c_double), allocatable, target :: y_zwt(:)
  real(c_double), allocatable, target :: y_zwt_perched(:)
  real(c_double), allocatable, target :: y_frost_table(:)
  real(c_double), allocatable, target :: y_wtfact(:)
  real(c_double), allocatable, target :: y_topo_slope(:)
  real(c_double), allocatable, target :: y_h2osfc_thresh(:)
  real(c_double), allocatable, target :: y_frac_h2osfc_nosnow(:)
  real(c_double), allocatable, target :: y_fsat(:)
  real(c_double), allocatable, target :: y_fcov(:)
  real(c_double), allocatable, target :: y_qflx_sat_excess_surf(:)
  real(c_double), allocatable, target :: y_qflx_top_soil(:)
  real(c_double), allocatable, target :: y_qflx_in_soil(:)
  real(c_double), allocatable, target :: y_qflx_top_soil_to_h2osfc(:)
  real(c_double), allocatable, target :: y_qinmax(:)
  real(c_double), allocatable, target :: y_qflx_infl_excess(:)
  real(c_double), allocatable, target :: y_qflx_in_soil_limited(:)
  real(c_double), allocatable, target :: y_qflx_in_h2osfc(:)
  real(c_double), allocatable, target :: y_qflx_infl_excess_surf(:)
  real(c_double), allocatable, target :: y_qflx_h2osfc_surf(:)
  real(c_double), allocatable, target :: y_qflx_h2osfc_drain(:)
  real(c_double), allocatable, target :: y_qflx_surf(:)
  real(c_double), allocatable, target :: y_h2osoi_vol(:,:)
  real(c_double), allocatable, target :: y_qflx_liqdew_to_top_layer(:)
  real(c_double), allocatable, target :: y_qflx_soliddew_to_top_layer(:)
  real(c_double), allocatable, target :: y_qflx_solidevap_from_top_layer(:)
  real(c_double), allocatable, target :: y_qflx_snwcp_liq(:)
  real(c_double), allocatable, target :: y_qflx_drain_perched(:)
  real(c_double), allocatable, target :: y_qflx_latflow_out(:)
  real(c_double), allocatable, target :: y_qflx_latflow_in(:)
  real(c_double), allocatable, target :: y_qflx_rsub_sat(:)
  real(c_double), allocatable, target :: y_qflx_ice_runoff_xs(:)
  real(c_double), allocatable, target :: y_qflx_drain(:)
  real(c_double), allocatable, target :: y_qflx_qrgwl(:)
  real(c_double), allocatable, target :: y_qflx_runoff(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, npatch, ngrid, nlevsno, nlevsoi, nlevgrnd, mode, npft
  read(u) cfg
  read(u) prm
  allocate(pftbuf(npft, 22), pftibuf(npft, 2))
  read(u) pftbuf
  read(u) pftibuf
  read(u) hprm
  begc = 1
  endc = ncol
  begp = 1
  endp = npatch
  begg = 1
  endg = ngrid
  allocate(c_snl(begc:endc))
  allocate(c_nbedrock(begc:endc))
  allocate(c_patchi(begc:endc))
  allocate(c_patchf(begc:endc))
  allocate(c_dz(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_z(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_zi(begc:endc, -nlevsno:nlevgrnd))
  allocate(c_watsat(begc:endc, 1:nlevgrnd))
  allocate(c_bsw(begc:endc, 1:nlevgrnd))
  allocate(c_sucsat(begc:endc, 1:nlevgrnd))
  allocate(c_hksat(begc:endc, 1:nlevgrnd))
  allocate(c_tkmg(begc:endc, 1:nlevgrnd))
  allocate(c_tkdry(begc:endc, 1:nlevgrnd))
  allocate(c_csol(begc:endc, 1:nlevgrnd))
  allocate(c_eff_porosity(begc:endc, 1:nlevgrnd))
  allocate(c_smp_l(begc:endc, 1:nlevgrnd))
  allocate(c_hk_l(begc:endc, 1:nlevgrnd))
  allocate(c_thk(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_icefrac(begc:endc, 1:nlevgrnd))
  allocate(c_qcharge(begc:endc))
  allocate(c_num_substeps(begc:endc))
  allocate(c_qin(begc:endc, 1:nlevgrnd))
  allocate(c_qout(begc:endc, 1:nlevgrnd))
  allocate(c_h2osoi_liq(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_h2osoi_ice(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_h2osfc(begc:endc))
  allocate(c_h2osno_no_layers(begc:endc))
  allocate(c_int_snow(begc:endc))
  allocate(c_frac_sno_eff(begc:endc))
  allocate(c_frac_h2osfc(begc:endc))
  allocate(c_snow_depth(begc:endc))
  allocate(c_bw(begc:endc, -nlevsno+1:0))
  allocate(c_snomelt_accum(begc:endc))
  allocate(c_qflx_infl(begc:endc))
  allocate(c_qflx_rootsoi(begc:endc, 1:nlevgrnd))
  allocate(c_qflx_snomelt(begc:endc))
  allocate(c_qflx_snofrz(begc:endc))
  allocate(c_qflx_snow_drain(begc:endc))
  allocate(c_qflx_snomelt_lyr(begc:endc, -nlevsno+1:0))
  allocate(c_qflx_snofrz_lyr(begc:endc, -nlevsno+1:0))
  allocate(c_qflx_h2osfc_to_ice(begc:endc))
  allocate(c_t_soisno(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_t_h2osfc(begc:endc))
  allocate(c_t_grnd(begc:endc))
  allocate(c_fact(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_xmf(begc:endc))
  allocate(c_xmf_h2osfc(begc:endc))
  allocate(c_c_h2osfc(begc:endc))
  allocate(c_imelt(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(c_emg(begc:endc))
  allocate(c_htvp(begc:endc))
  allocate(c_eflx_bot(begc:endc))
  allocate(c_eflx_fgr12(begc:endc))
  allocate(c_eflx_fgr(begc:endc, 1:nlevgrnd))
  allocate(c_eflx_snomelt(begc:endc))
  allocate(c_eflx_h2osfc_to_snow(begc:endc))
  allocate(c_forc_lwrad(begc:endc))
  allocate(c_zwt(begc:endc))
  allocate(c_smpmin(begc:endc))
  allocate(p_column(begp:endp))
  allocate(p_wtcol(begp:endp))
  allocate(p_frac_veg_nosno(begp:endp))
  allocate(p_sabg(begp:endp))
  allocate(p_sabg_soil(begp:endp))
  allocate(p_sabg_snow(begp:endp))
  allocate(p_sabg_lyr(begp:endp, -nlevsno+1:1))
  allocate(p_dlrad(begp:endp))
  allocate(p_cgrnd(begp:endp))
  allocate(p_eflx_sh_grnd(begp:endp))
  allocate(p_eflx_sh_snow(begp:endp))
  allocate(p_eflx_sh_soil(begp:endp))
  allocate(p_eflx_sh_h2osfc(begp:endp))
  allocate(p_qflx_evap_soi(begp:endp))
  allocate(p_qflx_ev_snow(begp:endp))
  allocate(p_qflx_ev_soil(begp:endp))
  allocate(p_qflx_ev_h2osfc(begp:endp))
  allocate(p_eflx_gnet(begp:endp))
  allocate(p_dgnetdT(begp:endp))
  allocate(p_sabg_chk(begp:endp))
  allocate(g_itype(begp:endp))
  allocate(g_gridcell(begp:endp))
  allocate(g_active(begp:endp))
  allocate(g_forc_u(begg:endg))
  allocate(g_forc_v(begg:endg))
  allocate(g_forc_pco2(begg:endg))
  allocate(g_forc_po2(begg:endg))
  allocate(g_forc_hgt_u(begg:endg))
  allocate(g_forc_hgt_t(begg:endg))
  allocate(g_forc_hgt_q(begg:endg))
  allocate(g_dayl(begg:endg))
  allocate(g_max_dayl(begg:endg))
  allocate(g_near_local_noon(begg:endg))
  allocate(g_forc_q(begc:endc))
  allocate(g_forc_pbot(begc:endc))
  allocate(g_forc_th(begc:endc))
  allocate(g_forc_rho(begc:endc))
  allocate(g_thv(begc:endc))
  allocate(g_qg(begc:endc))
  allocate(g_qg_snow(begc:endc))
  allocate(g_qg_soil(begc:endc))
  allocate(g_qg_h2osfc(begc:endc))
  allocate(g_dqgdT(begc:endc))
  allocate(g_z0mg(begc:endc))
  allocate(g_soilbeta(begc:endc))
  allocate(g_soilresis(begc:endc))
  allocate(g_h2osoi_liqvol(begc:endc, 1:nlevgrnd))
  allocate(g_qflx_tran_veg_col(begc:endc))
  allocate(g_rootr_col(begc:endc, 1:nlevgrnd))
  allocate(g_elai(begp:endp))
  allocate(g_esai(begp:endp))
  allocate(g_laisun(begp:endp))
  allocate(g_laisha(begp:endp))
  allocate(g_htop(begp:endp))
  allocate(g_sabv(begp:endp))
  allocate(g_parsun_z(begp:endp))
  allocate(g_parsha_z(begp:endp))
  allocate(g_laisun_z(begp:endp))
  allocate(g_laisha_z(begp:endp))
  allocate(g_vcmaxcintsun(begp:endp))
  allocate(g_vcmaxcintsha(begp:endp))
  allocate(g_nrad(begp:endp))
  allocate(g_thm(begp:endp))
  allocate(g_emv(begp:endp))
  allocate(g_fwet(begp:endp))
  allocate(g_fdry(begp:endp))
  allocate(g_t_a10(begp:endp))
  allocate(g_o3coefvsun(begp:endp))
  allocate(g_o3coefvsha(begp:endp))
  allocate(g_o3coefgsun(begp:endp))
  allocate(g_o3coefgsha(begp:endp))
  allocate(g_rootfr(begp:endp, 1:nlevgrnd))
  allocate(g_displa(begp:endp))
  allocate(g_z0mv(begp:endp))
  allocate(g_t_veg(begp:endp))
  allocate(g_t_stem(begp:endp))
  allocate(g_snocan(begp:endp))
  allocate(g_liqcan(begp:endp))
  allocate(g_cgrnds(begp:endp))
  allocate(g_cgrndl(begp:endp))
  allocate(g_z0hv(begp:endp))
  allocate(g_z0qv(begp:endp))
  allocate(g_forc_hgt_u_patch(begp:endp))
  allocate(g_forc_hgt_t_patch(begp:endp))
  allocate(g_forc_hgt_q_patch(begp:endp))
  allocate(g_ram1(begp:endp))
  allocate(g_rb1(begp:endp))
  allocate(g_rah1(begp:endp))
  allocate(g_rah2(begp:endp))
  allocate(g_raw1(begp:endp))
  allocate(g_raw2(begp:endp))
  allocate(g_ustar(begp:endp))
  allocate(g_um(begp:endp))
  allocate(g_uaf(begp:endp))
  allocate(g_taf(begp:endp))
  allocate(g_qaf(begp:endp))
  allocate(g_obu(begp:endp))
  allocate(g_zeta(begp:endp))
  allocate(g_vpd(begp:endp))
  allocate(g_num_iter(begp:endp))
  allocate(g_u10(begp:endp))
  allocate(g_u10_clm(begp:endp))
  allocate(g_va(begp:endp))
  allocate(g_fv(begp:endp))
  allocate(g_vds(begp:endp))
  allocate(g_btran(begp:endp))
  allocate(g_rresis(begp:endp, 1:nlevgrnd))
  allocate(g_rootr(begp:endp, 1:nlevgrnd))
  allocate(g_eflx_sh_veg(begp:endp))
  allocate(g_eflx_sh_stem(begp:endp))
  allocate(g_taux(begp:endp))
  allocate(g_tauy(begp:endp))
  allocate(g_ulrad(begp:endp))
  allocate(g_dhsdt_canopy(begp:endp))
  allocate(g_qflx_tran_veg(begp:endp))
  allocate(g_qflx_evap_veg(begp:endp))
  allocate(g_rh_af(begp:endp))
  allocate(g_q_ref2m(begp:endp))
  allocate(g_rh_ref2m(begp:endp))
  allocate(g_rh_ref2m_r(begp:endp))
  allocate(g_vpd_ref2m(begp:endp))
  allocate(g_iwue_ln(begp:endp))
  allocate(g_t_ref2m(begp:endp))
  allocate(g_t_ref2m_r(begp:endp))
  allocate(g_t_skin(begp:endp))
  allocate(g_stem_biomass(begp:endp))
  allocate(g_leaf_biomass(begp:endp))
  allocate(g_dleaf_patch(begp:endp))
  allocate(g_rssun(begp:endp))
  allocate(g_rssha(begp:endp))
  allocate(g_psnsun(begp:endp))
  allocate(g_psnsha(begp:endp))
  allocate(g_psnsun_wc(begp:endp))
  allocate(g_psnsun_wj(begp:endp))
  allocate(g_psnsun_wp(begp:endp))
  allocate(g_psnsha_wc(begp:endp))
  allocate(g_psnsha_wj(begp:endp))
  allocate(g_psnsha_wp(begp:endp))
  allocate(g_fpsn(begp:endp))
  allocate(g_fpsn_wc(begp:endp))
  allocate(g_fpsn_wj(begp:endp))
  allocate(g_fpsn_wp(begp:endp))
  allocate(g_lmrsun(begp:endp))
  allocate(g_lmrsha(begp:endp))
  allocate(g_gs_mol_sun(begp:endp))
  allocate(g_gs_mol_sha(begp:endp))
  allocate(g_gs_mol_sun_ln(begp:endp))
  allocate(g_gs_mol_sha_ln(begp:endp))
  allocate(g_cisun_z(begp:endp))
  allocate(g_cisha_z(begp:endp))
  allocate(g_rssun_z(begp:endp))
  allocate(g_rssha_z(begp:endp))
  allocate(g_lmrsun_z(begp:endp))
  allocate(g_lmrsha_z(begp:endp))
  allocate(g_psnsun_z(begp:endp))
  allocate(g_psnsha_z(begp:endp))
  allocate(g_vcmax_z(begp:endp))
  allocate(g_tpu_z(begp:endp))
  allocate(g_kp_z(begp:endp))
  allocate(g_ac(begp:endp))
  allocate(g_aj(begp:endp))
  allocate(g_ap(begp:endp))
  allocate(g_ag(begp:endp))
  allocate(g_an(begp:endp))
  allocate(g_gb_mol(begp:endp))
  allocate(g_gs_mol(begp:endp))
  allocate(g_cp(begp:endp))
  allocate(g_kc(begp:endp))
  allocate(g_ko(begp:endp))
  allocate(g_qe(begp:endp))
  allocate(g_bbb(begp:endp))
  allocate(g_mbb(begp:endp))
  allocate(g_rh_leaf(begp:endp))
  allocate(g_vpd_can(begp:endp))
  allocate(g_lnca(begp:endp))
  allocate(g_c3flag(begp:endp))
  allocate(g_canopy_err(begp:endp))
  allocate(y_qflx_rain_plus_snomelt(begc:endc))
  allocate(y_qflx_snow_h2osfc(begc:endc))
  allocate(y_qflx_floodc(begc:endc))
  allocate(y_qflx_liqevap_from_top_layer(begc:endc))
  allocate(y_qflx_ev_soil(begc:endc))
  allocate(y_qflx_ev_h2osfc(begc:endc))
  allocate(y_zwt(begc:endc))
  allocate(y_zwt_perched(begc:endc))
  allocate(y_frost_table(begc:endc))
  allocate(y_wtfact(begc:endc))
  allocate(y_topo_slope(begc:endc))
  allocate(y_h2osfc_thresh(begc:endc))
  allocate(y_frac_h2osfc_nosnow(begc:endc))
  allocate(y_fsat(begc:endc))
  allocate(y_fcov(begc:endc))
  allocate(y_qflx_sat_excess_surf(begc:endc))
  allocate(y_qflx_top_soil(begc:endc))
  allocate(y_qflx_in_soil(begc:endc))
  allocate(y_qflx_top_soil_to_h2osfc(begc:endc))
  allocate(y_qinmax(begc:endc))
  allocate(y_qflx_infl_excess(begc:endc))
  allocate(y_qflx_in_soil_limited(begc:endc))
  allocate(y_qflx_in_h2osfc(begc:endc))
  allocate(y_qflx_infl_excess_surf(begc:endc))
  allocate(y_qflx_h2osfc_surf(begc:endc))
  allocate(y_qflx_h2osfc_drain(begc:endc))
  allocate(y_qflx_surf(begc:endc))
  allocate(y_h2osoi_vol(begc:endc, 1:nlevgrnd))
  allocate(y_qflx_liqdew_to_top_layer(begc:endc))
  allocate(y_qflx_soliddew_to_top_layer(begc:endc))
  allocate(y_qflx_solidevap_from_top_layer(begc:endc))
  allocate(y_qflx_snwcp_liq(begc:endc))
  allocate(y_qflx_drain_perched(begc:endc))
  allocate(y_qflx_latflow_out(begc:endc))
  allocate(y_qflx_latflow_in(begc:endc))
  allocate(y_qflx_rsub_sat(begc:endc))
  allocate(y_qflx_ice_runoff_xs(begc:endc))
  allocate(y_qflx_drain(begc:endc))
  allocate(y_qflx_qrgwl(begc:endc))
  allocate(y_qflx_runoff(begc:endc))
  read(u) c_snl
  read(u) c_nbedrock
  read(u) c_patchi
  read(u) c_patchf
  read(u) c_dz
  read(u) c_z
  read(u) c_zi
  read(u) c_watsat
  read(u) c_bsw
  read(u) c_sucsat
  read(u) c_hksat
  read(u) c_tkmg
  read(u) c_tkdry
  read(u) c_csol
  read(u) c_eff_porosity
  read(u) c_smp_l
  read(u) c_hk_l
  read(u) c_thk
  read(u) c_icefrac
  read(u) c_qcharge
  read(u) c_num_substeps
  read(u) c_qin
  read(u) c_qout
  read(u) c_h2osoi_liq
  read(u) c_h2osoi_ice
  read(u) c_h2osfc
  read(u) c_h2osno_no_layers
  read(u) c_int_snow
  read(u) c_frac_sno_eff
  read(u) c_frac_h2osfc
  read(u) c_snow_depth
  read(u) c_bw
  read(u) c_snomelt_accum
  read(u) c_qflx_infl
  read(u) c_qflx_rootsoi
  read(u) c_qflx_snomelt
  read(u) c_qflx_snofrz
  read(u) c_qflx_snow_drain
  read(u) c_qflx_snomelt_lyr
  read(u) c_qflx_snofrz_lyr
  read(u) c_qflx_h2osfc_to_ice
  read(u) c_t_soisno
  read(u) c_t_h2osfc
  read(u) c_t_grnd
  read(u) c_fact
  read(u) c_xmf
  read(u) c_xmf_h2osfc
  read(u) c_c_h2osfc
  read(u) c_imelt
  read(u) c_emg
  read(u) c_htvp
  read(u) c_eflx_bot
  read(u) c_eflx_fgr12
  read(u) c_eflx_fgr
  read(u) c_eflx_snomelt
  read(u) c_eflx_h2osfc_to_snow
  read(u) c_forc_lwrad
  read(u) c_zwt
  read(u) c_smpmin
  read(u) p_column
  read(u) p_wtcol
  read(u) p_frac_veg_nosno
  read(u) p_sabg
  read(u) p_sabg_soil
  read(u) p_sabg_snow
  read(u) p_sabg_lyr
  read(u) p_dlrad
  read(u) p_cgrnd
  read(u) p_eflx_sh_grnd
  read(u) p_eflx_sh_snow
  read(u) p_eflx_sh_soil
  read(u) p_eflx_sh_h2osfc
  read(u) p_qflx_evap_soi
  read(u) p_qflx_ev_snow
  read(u) p_qflx_ev_soil
  read(u) p_qflx_ev_h2osfc
  read(u) p_eflx_gnet
  read(u) p_dgnetdT
  read(u) p_sabg_chk
  read(u) g_itype
  read(u) g_gridcell
  read(u) g_active
  read(u) g_forc_u
  read(u) g_forc_v
  read(u) g_forc_pco2
  read(u) g_forc_po2
  read(u) g_forc_hgt_u
  read(u) g_forc_hgt_t
  read(u) g_forc_hgt_q
  read(u) g_dayl
  read(u) g_max_dayl
  read(u) g_near_local_noon
  read(u) g_forc_q
  read(u) g_forc_pbot
  read(u) g_forc_th
  read(u) g_forc_rho
  read(u) g_thv
  read(u) g_qg
  read(u) g_qg_snow
  read(u) g_qg_soil
  read(u) g_qg_h2osfc
  read(u) g_dqgdT
  read(u) g_z0mg
  read(u) g_soilbeta
  read(u) g_soilresis
  read(u) g_h2osoi_liqvol
  read(u) g_qflx_tran_veg_col
  read(u) g_rootr_col
  read(u) g_elai
  read(u) g_esai
  read(u) g_laisun
  read(u) g_laisha
  read(u) g_htop
  read(u) g_sabv
  read(u) g_parsun_z
  read(u) g_parsha_z
  read(u) g_laisun_z
  read(u) g_laisha_z
  read(u) g_vcmaxcintsun
  read(u) g_vcmaxcintsha
  read(u) g_nrad
  read(u) g_thm
  read(u) g_emv
  read(u) g_fwet
  read(u) g_fdry
  read(u) g_t_a10
  read(u) g_o3coefvsun
  read(u) g_o3coefvsha
  read(u) g_o3coefgsun
  read(u) g_o3coefgsha
  read(u) g_rootfr
  read(u) g_displa
  read(u) g_z0mv
  read(u) g_t_veg
  read(u) g_t_stem
  read(u) g_snocan
  read(u) g_liqcan
  read(u) g_cgrnds
  read(u) g_cgrndl
  read(u) g_z0hv
  read(u) g_z0qv
  read(u) g_forc_hgt_u_patch
  read(u) g_forc_hgt_t_patch
  read(u) g_forc_hgt_q_patch
  read(u) g_ram1
  read(u) g_rb1
  read(u) g_rah1
  read(u) g_rah2
  read(u) g_raw1
  read(u) g_raw2
  read(u) g_ustar
  read(u) g_um
  read(u) g_uaf
  read(u) g_taf
  read(u) g_qaf
  read(u) g_obu
  read(u) g_zeta
  read(u) g_vpd
  read(u) g_num_iter
  read(u) g_u10
  read(u) g_u10_clm
  read(u) g_va
  read(u) g_fv
  read(u) g_vds
  read(u) g_btran
  read(u) g_rresis
  read(u) g_rootr
  read(u) g_eflx_sh_veg
  read(u) g_eflx_sh_stem
  read(u) g_taux
  read(u) g_tauy
  read(u) g_ulrad
  read(u) g_dhsdt_canopy
  read(u) g_qflx_tran_veg
  read(u) g_qflx_evap_veg
  read(u) g_rh_af
  read(u) g_q_ref2m
  read(u) g_rh_ref2m
  read(u) g_rh_ref2m_r
  read(u) g_vpd_ref2m
  read(u) g_iwue_ln
  read(u) g_t_ref2m
  read(u) g_t_ref2m_r
  read(u) g_t_skin
  read(u) g_stem_biomass
  read(u) g_leaf_biomass
  read(u) g_dleaf_patch
  read(u) g_rssun
  read(u) g_rssha
  read(u) g_psnsun
  read(u) g_psnsha
  read(u) g_psnsun_wc
  read(u) g_psnsun_wj
  read(u) g_psnsun_wp
  read(u) g_psnsha_wc
  read(u) g_psnsha_wj
  read(u) g_psnsha_wp
  read(u) g_fpsn
  read(u) g_fpsn_wc
  read(u) g_fpsn_wj
  read(u) g_fpsn_wp
  read(u) g_lmrsun
  read(u) g_lmrsha
  read(u) g_gs_mol_sun
  read(u) g_gs_mol_sha
  read(u) g_gs_mol_sun_ln
  read(u) g_gs_mol_sha_ln
  read(u) g_cisun_z
  read(u) g_cisha_z
  read(u) g_rssun_z
  read(u) g_rssha_z
  read(u) g_lmrsun_z
  read(u) g_lmrsha_z
  read(u) g_psnsun_z
  read(u) g_psnsha_z
  read(u) g_vcmax_z
  read(u) g_tpu_z
  read(u) g_kp_z
  read(u) g_ac
  read(u) g_aj
  read(u) g_ap
  read(u) g_ag
  read(u) g_an
  read(u) g_gb_mol
  read(u) g_gs_mol
  read(u) g_cp
  read(u) g_kc
  read(u) g_ko
  read(u) g_qe
  read(u) g_bbb
  read(u) g_mbb
  read(u) g_rh_leaf
  read(u) g_vpd_can
  read(u) g_lnca
  read(u) g_c3flag
  read(u) g_canopy_err
  read(u) y_qflx_rain_plus_snomelt
  read(u) y_qflx_snow_h2osfc
  read(u) y_qflx_floodc
  read(u) y_qflx_liqevap_from_top_layer
  read(u) y_qflx_ev_soil
  read(u) y_qflx_ev_h2osfc
  read(u) y_zwt
  read(u) y_zwt_perched
  read(u) y_frost_table
  read(u) y_wtfact
  read(u) y_topo_slope
  read(u) y_h2osfc_thresh
  read(u) y_frac_h2osfc_nosnow
  read(u) y_fsat
  read(u) y_fcov
  read(u) y_qflx_sat_excess_surf
  read(u) y_qflx_top_soil
  read(u) y_qflx_in_soil
  read(u) y_qflx_top_soil_to_h2osfc
  read(u) y_qinmax
  read(u) y_qflx_infl_excess
  read(u) y_qflx_in_soil_limited
  read(u) y_qflx_in_h2osfc
  read(u) y_qflx_infl_excess_surf
  read(u) y_qflx_h2osfc_surf
  read(u) y_qflx_h2osfc_drain
  read(u) y_qflx_surf
  read(u) y_h2osoi_vol
  read(u) y_qflx_liqdew_to_top_layer
  read(u) y_qflx_soliddew_to_top_layer
  read(u) y_qflx_solidevap_from_top_layer
  read(u) y_qflx_snwcp_liq
  read(u) y_qflx_drain_perched
  read(u) y_qflx_latflow_out
  read(u) y_qflx_latflow_in
  read(u) y_qflx_rsub_sat
  read(u) y_qflx_ice_runoff_xs
  read(u) y_qflx_drain
  read(u) y_qflx_qrgwl
  read(u) y_qflx_runoff
  close(u)

  col%snl = c_loc(c_snl(begc))
  col%nbedrock = c_loc(c_nbedrock(begc))
  col%patchi = c_loc(c_patchi(begc))
  col%patchf = c_loc(c_patchf(begc))
  col%dz = c_loc(c_dz(begc, -nlevsno+1))
  col%z = c_loc(c_z(begc, -nlevsno+1))
  col%zi = c_loc(c_zi(begc, -nlevsno))
  col%watsat = c_loc(c_watsat(begc, 1))
  col%bsw = c_loc(c_bsw(begc, 1))
  col%sucsat = c_loc(c_sucsat(begc, 1))
  col%hksat = c_loc(c_hksat(begc, 1))
  col%tkmg = c_loc(c_tkmg(begc, 1))
  col%tkdry = c_loc(c_tkdry(begc, 1))
  col%csol = c_loc(c_csol(begc, 1))
  col%eff_porosity = c_loc(c_eff_porosity(begc, 1))
  col%smp_l = c_loc(c_smp_l(begc, 1))
  col%hk_l = c_loc(c_hk_l(begc, 1))
  col%thk = c_loc(c_thk(begc, -nlevsno+1))
  col%icefrac = c_loc(c_icefrac(begc, 1))
  col%qcharge = c_loc(c_qcharge(begc))
  col%num_substeps = c_loc(c_num_substeps(begc))
  col%qin = c_loc(c_qin(begc, 1))
  col%qout = c_loc(c_qout(begc, 1))
  col%h2osoi_liq = c_loc(c_h2osoi_liq(begc, -nlevsno+1))
  col%h2osoi_ice = c_loc(c_h2osoi_ice(begc, -nlevsno+1))
  col%h2osfc = c_loc(c_h2osfc(begc))
  col%h2osno_no_layers = c_loc(c_h2osno_no_layers(begc))
  col%int_snow = c_loc(c_int_snow(begc))
  col%frac_sno_eff = c_loc(c_frac_sno_eff(begc))
  col%frac_h2osfc = c_loc(c_frac_h2osfc(begc))
  col%snow_depth = c_loc(c_snow_depth(begc))
  col%bw = c_loc(c_bw(begc, -nlevsno+1))
  col%snomelt_accum = c_loc(c_snomelt_accum(begc))
  col%qflx_infl = c_loc(c_qflx_infl(begc))
  col%qflx_rootsoi = c_loc(c_qflx_rootsoi(begc, 1))
  col%qflx_snomelt = c_loc(c_qflx_snomelt(begc))
  col%qflx_snofrz = c_loc(c_qflx_snofrz(begc))
  col%qflx_snow_drain = c_loc(c_qflx_snow_drain(begc))
  col%qflx_snomelt_lyr = c_loc(c_qflx_snomelt_lyr(begc, -nlevsno+1))
  col%qflx_snofrz_lyr = c_loc(c_qflx_snofrz_lyr(begc, -nlevsno+1))
  col%qflx_h2osfc_to_ice = c_loc(c_qflx_h2osfc_to_ice(begc))
  col%t_soisno = c_loc(c_t_soisno(begc, -nlevsno+1))
  col%t_h2osfc = c_loc(c_t_h2osfc(begc))
  col%t_grnd = c_loc(c_t_grnd(begc))
  col%fact = c_loc(c_fact(begc, -nlevsno+1))
  col%xmf = c_loc(c_xmf(begc))
  col%xmf_h2osfc = c_loc(c_xmf_h2osfc(begc))
  col%c_h2osfc = c_loc(c_c_h2osfc(begc))
  col%imelt = c_loc(c_imelt(begc, -nlevsno+1))
  col%emg = c_loc(c_emg(begc))
  col%htvp = c_loc(c_htvp(begc))
  col%eflx_bot = c_loc(c_eflx_bot(begc))
  col%eflx_fgr12 = c_loc(c_eflx_fgr12(begc))
  col%eflx_fgr = c_loc(c_eflx_fgr(begc, 1))
  col%eflx_snomelt = c_loc(c_eflx_snomelt(begc))
  col%eflx_h2osfc_to_snow = c_loc(c_eflx_h2osfc_to_snow(begc))
  col%forc_lwrad = c_loc(c_forc_lwrad(begc))
  col%zwt = c_loc(c_zwt(begc))
  col%smpmin = c_loc(c_smpmin(begc))
  pch%column = c_loc(p_column(begp))
  pch%wtcol = c_loc(p_wtcol(begp))
  pch%frac_veg_nosno = c_loc(p_frac_veg_nosno(begp))
  pch%sabg = c_loc(p_sabg(begp))
  pch%sabg_soil = c_loc(p_sabg_soil(begp))
  pch%sabg_snow = c_loc(p_sabg_snow(begp))
  pch%sabg_lyr = c_loc(p_sabg_lyr(begp, -nlevsno+1))
  pch%dlrad = c_loc(p_dlrad(begp))
  pch%cgrnd = c_loc(p_cgrnd(begp))
  pch%eflx_sh_grnd = c_loc(p_eflx_sh_grnd(begp))
  pch%eflx_sh_snow = c_loc(p_eflx_sh_snow(begp))
  pch%eflx_sh_soil = c_loc(p_eflx_sh_soil(begp))
  pch%eflx_sh_h2osfc = c_loc(p_eflx_sh_h2osfc(begp))
  pch%qflx_evap_soi = c_loc(p_qflx_evap_soi(begp))
  pch%qflx_ev_snow = c_loc(p_qflx_ev_snow(begp))
  pch%qflx_ev_soil = c_loc(p_qflx_ev_soil(begp))
  pch%qflx_ev_h2osfc = c_loc(p_qflx_ev_h2osfc(begp))
  pch%eflx_gnet = c_loc(p_eflx_gnet(begp))
  pch%dgnetdT = c_loc(p_dgnetdT(begp))
  pch%sabg_chk = c_loc(p_sabg_chk(begp))
  can%itype = c_loc(g_itype(begp))
  can%gridcell = c_loc(g_gridcell(begp))
  can%active = c_loc(g_active(begp))
  can%forc_u = c_loc(g_forc_u(begg))
  can%forc_v = c_loc(g_forc_v(begg))
  can%forc_pco2 = c_loc(g_forc_pco2(begg))
  can%forc_po2 = c_loc(g_forc_po2(begg))
  can%forc_hgt_u = c_loc(g_forc_hgt_u(begg))
  can%forc_hgt_t = c_loc(g_forc_hgt_t(begg))
  can%forc_hgt_q = c_loc(g_forc_hgt_q(begg))
  can%dayl = c_loc(g_dayl(begg))
  can%max_dayl = c_loc(g_max_dayl(begg))
  can%near_local_noon = c_loc(g_near_local_noon(begg))
  can%forc_q = c_loc(g_forc_q(begc))
  can%forc_pbot = c_loc(g_forc_pbot(begc))
  can%forc_th = c_loc(g_forc_th(begc))
  can%forc_rho = c_loc(g_forc_rho(begc))
  can%thv = c_loc(g_thv(begc))
  can%qg = c_loc(g_qg(begc))
  can%qg_snow = c_loc(g_qg_snow(begc))
  can%qg_soil = c_loc(g_qg_soil(begc))
  can%qg_h2osfc = c_loc(g_qg_h2osfc(begc))
  can%dqgdT = c_loc(g_dqgdT(begc))
  can%z0mg = c_loc(g_z0mg(begc))
  can%soilbeta = c_loc(g_soilbeta(begc))
  can%soilresis = c_loc(g_soilresis(begc))
  can%h2osoi_liqvol = c_loc(g_h2osoi_liqvol(begc, 1))
  can%qflx_tran_veg_col = c_loc(g_qflx_tran_veg_col(begc))
  can%rootr_col = c_loc(g_rootr_col(begc, 1))
  can%elai = c_loc(g_elai(begp))
  can%esai = c_loc(g_esai(begp))
  can%laisun = c_loc(g_laisun(begp))
  can%laisha = c_loc(g_laisha(begp))
  can%htop = c_loc(g_htop(begp))
  can%sabv = c_loc(g_sabv(begp))
  can%parsun_z = c_loc(g_parsun_z(begp))
  can%parsha_z = c_loc(g_parsha_z(begp))
  can%laisun_z = c_loc(g_laisun_z(begp))
  can%laisha_z = c_loc(g_laisha_z(begp))
  can%vcmaxcintsun = c_loc(g_vcmaxcintsun(begp))
  can%vcmaxcintsha = c_loc(g_vcmaxcintsha(begp))
  can%nrad = c_loc(g_nrad(begp))
  can%thm = c_loc(g_thm(begp))
  can%emv = c_loc(g_emv(begp))
  can%fwet = c_loc(g_fwet(begp))
  can%fdry = c_loc(g_fdry(begp))
  can%t_a10 = c_loc(g_t_a10(begp))
  can%o3coefvsun = c_loc(g_o3coefvsun(begp))
  can%o3coefvsha = c_loc(g_o3coefvsha(begp))
  can%o3coefgsun = c_loc(g_o3coefgsun(begp))
  can%o3coefgsha = c_loc(g_o3coefgsha(begp))
  can%rootfr = c_loc(g_rootfr(begp, 1))
  can%displa = c_loc(g_displa(begp))
  can%z0mv = c_loc(g_z0mv(begp))
  can%t_veg = c_loc(g_t_veg(begp))
  can%t_stem = c_loc(g_t_stem(begp))
  can%snocan = c_loc(g_snocan(begp))
  can%liqcan = c_loc(g_liqcan(begp))
  can%cgrnds = c_loc(g_cgrnds(begp))
  can%cgrndl = c_loc(g_cgrndl(begp))
  can%z0hv = c_loc(g_z0hv(begp))
  can%z0qv = c_loc(g_z0qv(begp))
  can%forc_hgt_u_patch = c_loc(g_forc_hgt_u_patch(begp))
  can%forc_hgt_t_patch = c_loc(g_forc_hgt_t_patch(begp))
  can%forc_hgt_q_patch = c_loc(g_forc_hgt_q_patch(begp))
  can%ram1 = c_loc(g_ram1(begp))
  can%rb1 = c_loc(g_rb1(begp))
  can%rah1 = c_loc(g_rah1(begp))
  can%rah2 = c_loc(g_rah2(begp))
  can%raw1 = c_loc(g_raw1(begp))
  can%raw2 = c_loc(g_raw2(begp))
  can%ustar = c_loc(g_ustar(begp))
  can%um = c_loc(g_um(begp))
  can%uaf = c_loc(g_uaf(begp))
  can%taf = c_loc(g_taf(begp))
  can%qaf = c_loc(g_qaf(begp))
  can%obu = c_loc(g_obu(begp))
  can%zeta = c_loc(g_zeta(begp))
  can%vpd = c_loc(g_vpd(begp))
  can%num_iter = c_loc(g_num_iter(begp))
  can%u10 = c_loc(g_u10(begp))
  can%u10_clm = c_loc(g_u10_clm(begp))
  can%va = c_loc(g_va(begp))
  can%fv = c_loc(g_fv(begp))
  can%vds = c_loc(g_vds(begp))
  can%btran = c_loc(g_btran(begp))
  can%rresis = c_loc(g_rresis(begp, 1))
  can%rootr = c_loc(g_rootr(begp, 1))
  can%eflx_sh_veg = c_loc(g_eflx_sh_veg(begp))
  can%eflx_sh_stem = c_loc(g_eflx_sh_stem(begp))
  can%taux = c_loc(g_taux(begp))
  can%tauy = c_loc(g_tauy(begp))
  can%ulrad = c_loc(g_ulrad(begp))
  can%dhsdt_canopy = c_loc(g_dhsdt_canopy(begp))
  can%qflx_tran_veg = c_loc(g_qflx_tran_veg(begp))
  can%qflx_evap_veg = c_loc(g_qflx_evap_veg(begp))
  can%rh_af = c_loc(g_rh_af(begp))
  can%q_ref2m = c_loc(g_q_ref2m(begp))
  can%rh_ref2m = c_loc(g_rh_ref2m(begp))
  can%rh_ref2m_r = c_loc(g_rh_ref2m_r(begp))
  can%vpd_ref2m = c_loc(g_vpd_ref2m(begp))
  can%iwue_ln = c_loc(g_iwue_ln(begp))
  can%t_ref2m = c_loc(g_t_ref2m(begp))
  can%t_ref2m_r = c_loc(g_t_ref2m_r(begp))
  can%t_skin = c_loc(g_t_skin(begp))
  can%stem_biomass = c_loc(g_stem_biomass(begp))
  can%leaf_biomass = c_loc(g_leaf_biomass(begp))
  can%dleaf_patch = c_loc(g_dleaf_patch(begp))
  can%rssun = c_loc(g_rssun(begp))
  can%rssha = c_loc(g_rssha(begp))
  can%psnsun = c_loc(g_psnsun(begp))
  can%psnsha = c_loc(g_psnsha(begp))
  can%psnsun_wc = c_loc(g_psnsun_wc(begp))
  can%psnsun_wj = c_loc(g_psnsun_wj(begp))
  can%psnsun_wp = c_loc(g_psnsun_wp(begp))
  can%psnsha_wc = c_loc(g_psnsha_wc(begp))
  can%psnsha_wj = c_loc(g_psnsha_wj(begp))
  can%psnsha_wp = c_loc(g_psnsha_wp(begp))
  can%fpsn = c_loc(g_fpsn(begp))
  can%fpsn_wc = c_loc(g_fpsn_wc(begp))
  can%fpsn_wj = c_loc(g_fpsn_wj(begp))
  can%fpsn_wp = c_loc(g_fpsn_wp(begp))
  can%lmrsun = c_loc(g_lmrsun(begp))
  can%lmrsha = c_loc(g_lmrsha(begp))
  can%gs_mol_sun = c_loc(g_gs_mol_sun(begp))
  can%gs_mol_sha = c_loc(g_gs_mol_sha(begp))
  can%gs_mol_sun_ln = c_loc(g_gs_mol_sun_ln(begp))
  can%gs_mol_sha_ln = c_loc(g_gs_mol_sha_ln(begp))
  can%cisun_z = c_loc(g_cisun_z(begp))
  can%cisha_z = c_loc(g_cisha_z(begp))
  can%rssun_z = c_loc(g_rssun_z(begp))
  can%rssha_z = c_loc(g_rssha_z(begp))
  can%lmrsun_z = c_loc(g_lmrsun_z(begp))
  can%lmrsha_z = c_loc(g_lmrsha_z(begp))
  can%psnsun_z = c_loc(g_psnsun_z(begp))
  can%psnsha_z = c_loc(g_psnsha_z(begp))
  can%vcmax_z = c_loc(g_vcmax_z(begp))
  can%tpu_z = c_loc(g_tpu_z(begp))
  can%kp_z = c_loc(g_kp_z(begp))
  can%ac = c_loc(g_ac(begp))
  can%aj = c_loc(g_aj(begp))
  can%ap = c_loc(g_ap(begp))
  can%ag = c_loc(g_ag(begp))
  can%an = c_loc(g_an(begp))
  can%gb_mol = c_loc(g_gb_mol(begp))
  can%gs_mol = c_loc(g_gs_mol(begp))
  can%cp = c_loc(g_cp(begp))
  can%kc = c_loc(g_kc(begp))
  can%ko = c_loc(g_ko(begp))
  can%qe = c_loc(g_qe(begp))
  can%bbb = c_loc(g_bbb(begp))
  can%mbb = c_loc(g_mbb(begp))
  can%rh_leaf = c_loc(g_rh_leaf(begp))
  can%vpd_can = c_loc(g_vpd_can(begp))
  can%lnca = c_loc(g_lnca(begp))
  can%c3flag = c_loc(g_c3flag(begp))
  can%canopy_err = c_loc(g_canopy_err(begp))
  can%column = pch%column
  can%wtcol = pch%wtcol
  can%frac_veg_nosno = pch%frac_veg_nosno
  can%dlrad = pch%dlrad
  can%cgrnd = pch%cgrnd
  can%eflx_sh_grnd = pch%eflx_sh_grnd
  can%eflx_sh_snow = pch%eflx_sh_snow
  can%eflx_sh_soil = pch%eflx_sh_soil
  can%eflx_sh_h2osfc = pch%eflx_sh_h2osfc
  can%qflx_evap_soi = pch%qflx_evap_soi
  can%qflx_ev_snow = pch%qflx_ev_snow
  can%qflx_ev_soil = pch%qflx_ev_soil
  can%qflx_ev_h2osfc = pch%qflx_ev_h2osfc
  can%qflx_rootsoi_col = col%qflx_rootsoi
  sw%qflx_rain_plus_snomelt = c_loc(y_qflx_rain_plus_snomelt(begc))
  sw%qflx_snow_h2osfc = c_loc(y_qflx_snow_h2osfc(begc))
  sw%qflx_floodc = c_loc(y_qflx_floodc(begc))
  sw%qflx_liqevap_from_top_layer = c_loc(y_qflx_liqevap_from_top_layer(begc))
  sw%qflx_ev_soil = c_loc(y_qflx_ev_soil(begc))
  sw%qflx_ev_h2osfc = c_loc(y_qflx_ev_h2osfc(begc))
  sw%zwt = c_loc(y_zwt(begc))
  sw%zwt_perched = c_loc(y_zwt_perched(begc))
  sw%frost_table = c_loc(y_frost_table(begc))
  sw%wtfact = c_loc(y_wtfact(begc))
  sw%topo_slope = c_loc(y_topo_slope(begc))
  sw%h2osfc_thresh = c_loc(y_h2osfc_thresh(begc))
  sw%frac_h2osfc_nosnow = c_loc(y_frac_h2osfc_nosnow(begc))
  sw%h2osfc = col%h2osfc
  sw%fsat = c_loc(y_fsat(begc))
  sw%fcov = c_loc(y_fcov(begc))
  sw%qflx_sat_excess_surf = c_loc(y_qflx_sat_excess_surf(begc))
  sw%qflx_top_soil = c_loc(y_qflx_top_soil(begc))
  sw%qflx_in_soil = c_loc(y_qflx_in_soil(begc))
  sw%qflx_top_soil_to_h2osfc = c_loc(y_qflx_top_soil_to_h2osfc(begc))
  sw%qinmax = c_loc(y_qinmax(begc))
  sw%qflx_infl_excess = c_loc(y_qflx_infl_excess(begc))
  sw%qflx_in_soil_limited = c_loc(y_qflx_in_soil_limited(begc))
  sw%qflx_in_h2osfc = c_loc(y_qflx_in_h2osfc(begc))
  sw%qflx_infl_excess_surf = c_loc(y_qflx_infl_excess_surf(begc))
  sw%qflx_h2osfc_surf = c_loc(y_qflx_h2osfc_surf(begc))
  sw%qflx_h2osfc_drain = c_loc(y_qflx_h2osfc_drain(begc))
  sw%qflx_infl = col%qflx_infl
  sw%qflx_surf = c_loc(y_qflx_surf(begc))
  wt%zwt = c_loc(y_zwt(begc))
  wt%zwt_perched = c_loc(y_zwt_perched(begc))
  wt%frost_table = c_loc(y_frost_table(begc))
  wt%h2osoi_vol = c_loc(y_h2osoi_vol(begc, 1))
  wt%qflx_liqdew_to_top_layer = c_loc(y_qflx_liqdew_to_top_layer(begc))
  wt%qflx_soliddew_to_top_layer = c_loc(y_qflx_soliddew_to_top_layer(begc))
  wt%qflx_solidevap_from_top_layer = c_loc(y_qflx_solidevap_from_top_layer(begc))
  dr%topo_slope = c_loc(y_topo_slope(begc))
  dr%zwt_perched = c_loc(y_zwt_perched(begc))
  dr%frost_table = c_loc(y_frost_table(begc))
  dr%qflx_surf = c_loc(y_qflx_surf(begc))
  dr%qflx_snwcp_liq = c_loc(y_qflx_snwcp_liq(begc))
  dr%zwt = c_loc(y_zwt(begc))
  dr%h2osoi_vol = c_loc(y_h2osoi_vol(begc, 1))
  dr%qflx_drain_perched = c_loc(y_qflx_drain_perched(begc))
  dr%qflx_latflow_out = c_loc(y_qflx_latflow_out(begc))
  dr%qflx_latflow_in = c_loc(y_qflx_latflow_in(begc))
  dr%qflx_rsub_sat = c_loc(y_qflx_rsub_sat(begc))
  dr%qflx_ice_runoff_xs = c_loc(y_qflx_ice_runoff_xs(begc))
  dr%qflx_drain = c_loc(y_qflx_drain(begc))
  dr%qflx_qrgwl = c_loc(y_qflx_qrgwl(begc))
  dr%qflx_runoff = c_loc(y_qflx_runoff(begc))
  pft%npft = npft
  pft%dleaf = c_loc(pftbuf(1, 1))
  pft%c3psn = c_loc(pftbuf(1, 2))
  pft%slatop = c_loc(pftbuf(1, 3))
  pft%leafcn = c_loc(pftbuf(1, 4))
  pft%flnr = c_loc(pftbuf(1, 5))
  pft%fnitr = c_loc(pftbuf(1, 6))
  pft%mbbopt = c_loc(pftbuf(1, 7))
  pft%medlynslope = c_loc(pftbuf(1, 8))
  pft%medlynintercept = c_loc(pftbuf(1, 9))
  pft%smpso = c_loc(pftbuf(1, 10))
  pft%smpsc = c_loc(pftbuf(1, 11))
  pft%theta_cj = c_loc(pftbuf(1, 12))
  pft%z0v_Cr = c_loc(pftbuf(1, 13))
  pft%z0v_Cs = c_loc(pftbuf(1, 14))
  pft%z0v_c = c_loc(pftbuf(1, 15))
  pft%z0v_cw = c_loc(pftbuf(1, 16))
  pft%z0v_LAImax = c_loc(pftbuf(1, 17))
  pft%dbh = c_loc(pftbuf(1, 18))
  pft%fbw = c_loc(pftbuf(1, 19))
  pft%nstem = c_loc(pftbuf(1, 20))
  pft%rstem_per_dbh = c_loc(pftbuf(1, 21))
  pft%wood_density = c_loc(pftbuf(1, 22))
  pft%is_tree = c_loc(pftibuf(1, 1))
  pft%is_shrub = c_loc(pftibuf(1, 2))

  rc = ctsm_create(cfg, 0, ctx)
  if (rc /= CTSM_OK) then
     print '(I0)', rc
     stop 1
  end if
  lay%begc_alloc = begc
  lay%ldc = endc - begc + 1
  lay%begp_alloc = begp
  lay%ldp = endp - begp + 1
  lay%begg_alloc = begg
  lay%ldg = endg - begg + 1
  select case (mode)
  case (1)
     allocate(filter(ncol))
     filter = [(i, i = begc, endc)]
     rc = ctsm_soilwater(ctx, lay, col, ncol, filter, 1800.0_c_double)
  case (2)
     allocate(filter(ncol))
     filter = [(i, i = begc, endc)]
     rc = ctsm_soiltemperature(ctx, lay, col, pch, ncol, filter, 1800.0_c_double)
  case (3)
     rc = ctsm_set_canopy_params(ctx, prm, pft)
     if (rc == CTSM_OK) then
        allocate(filter(npatch))
        filter = [(i, i = begp, endp)]
        rc = ctsm_canopyfluxes(ctx, lay, col, can, npatch, filter, 1800.0_c_double)
     end if
  case (4)
     allocate(filter(ncol))
     filter = [(i, i = begc, endc)]
     rc = ctsm_surface_water(ctx, lay, col, sw, hprm, ncol, filter, 1800.0_c_double)
  case (5)
     allocate(filter(ncol))
     filter = [(i, i = begc, endc)]
     rc = ctsm_water_table(ctx, lay, col, wt, ncol, filter, 1800.0_c_double)
  case (6)
     allocate(filter(ncol))
     filter = [(i, i = begc, endc)]
     rc = ctsm_drainage(ctx, lay, col, dr, hprm, ncol, filter, 1800.0_c_double)
  case default
     rc = CTSM_E_ARG
  end select
  print '(I0)', rc
  i = ctsm_destroy(ctx)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) c_snl
  write(u) c_nbedrock
  write(u) c_patchi
  write(u) c_patchf
  write(u) c_dz
  write(u) c_z
  write(u) c_zi
  write(u) c_watsat
  write(u) c_bsw
  write(u) c_sucsat
  write(u) c_hksat
  write(u) c_tkmg
  write(u) c_tkdry
  write(u) c_csol
  write(u) c_eff_porosity
  write(u) c_smp_l
  write(u) c_hk_l
  write(u) c_thk
  write(u) c_icefrac
  write(u) c_qcharge
  write(u) c_num_substeps
  write(u) c_qin
  write(u) c_qout
  write(u) c_h2osoi_liq
  write(u) c_h2osoi_ice
  write(u) c_h2osfc
  write(u) c_h2osno_no_layers
  write(u) c_int_snow
  write(u) c_frac_sno_eff
  write(u) c_frac_h2osfc
  write(u) c_snow_depth
  write(u) c_bw
  write(u) c_snomelt_accum
  write(u) c_qflx_infl
  write(u) c_qflx_rootsoi
  write(u) c_qflx_snomelt
  write(u) c_qflx_snofrz
  write(u) c_qflx_snow_drain
  write(u) c_qflx_snomelt_lyr
  write(u) c_qflx_snofrz_lyr
  write(u) c_qflx_h2osfc_to_ice
  write(u) c_t_soisno
  write(u) c_t_h2osfc
  write(u) c_t_grnd
  write(u) c_fact
  write(u) c_xmf
  write(u) c_xmf_h2osfc
  write(u) c_c_h2osfc
  write(u) c_imelt
  write(u) c_emg
  write(u) c_htvp
  write(u) c_eflx_bot
  write(u) c_eflx_fgr12
  write(u) c_eflx_fgr
  write(u) c_eflx_snomelt
  write(u) c_eflx_h2osfc_to_snow
  write(u) c_forc_lwrad
  write(u) c_zwt
  write(u) c_smpmin
  write(u) p_column
  write(u) p_wtcol
  write(u) p_frac_veg_nosno
  write(u) p_sabg
  write(u) p_sabg_soil
  write(u) p_sabg_snow
  write(u) p_sabg_lyr
  write(u) p_dlrad
  write(u) p_cgrnd
  write(u) p_eflx_sh_grnd
  write(u) p_eflx_sh_snow
  write(u) p_eflx_sh_soil
  write(u) p_eflx_sh_h2osfc
  write(u) p_qflx_evap_soi
  write(u) p_qflx_ev_snow
  write(u) p_qflx_ev_soil
  write(u) p_qflx_ev_h2osfc
  write(u) p_eflx_gnet
  write(u) p_dgnetdT
  write(u) p_sabg_chk
  write(u) g_itype
  write(u) g_gridcell
  write(u) g_active
  write(u) g_forc_u
  write(u) g_forc_v
  write(u) g_forc_pco2
  write(u) g_forc_po2
  write(u) g_forc_hgt_u
  write(u) g_forc_hgt_t
  write(u) g_forc_hgt_q
  write(u) g_dayl
  write(u) g_max_dayl
  write(u) g_near_local_noon
  write(u) g_forc_q
  write(u) g_forc_pbot
  write(u) g_forc_th
  write(u) g_forc_rho
  write(u) g_thv
  write(u) g_qg
  write(u) g_qg_snow
  write(u) g_qg_soil
  write(u) g_qg_h2osfc
  write(u) g_dqgdT
  write(u) g_z0mg
  write(u) g_soilbeta
  write(u) g_soilresis
  write(u) g_h2osoi_liqvol
  write(u) g_qflx_tran_veg_col
  write(u) g_rootr_col
  write(u) g_elai
  write(u) g_esai
  write(u) g_laisun
  write(u) g_laisha
  write(u) g_htop
  write(u) g_sabv
  write(u) g_parsun_z
  write(u) g_parsha_z
  write(u) g_laisun_z
  write(u) g_laisha_z
  write(u) g_vcmaxcintsun
  write(u) g_vcmaxcintsha
  write(u) g_nrad
  write(u) g_thm
  write(u) g_emv
  write(u) g_fwet
  write(u) g_fdry
  write(u) g_t_a10
  write(u) g_o3coefvsun
  write(u) g_o3coefvsha
  write(u) g_o3coefgsun
  write(u) g_o3coefgsha
  write(u) g_rootfr
  write(u) g_displa
  write(u) g_z0mv
  write(u) g_t_veg
  write(u) g_t_stem
  write(u) g_snocan
  write(u) g_liqcan
  write(u) g_cgrnds
  write(u) g_cgrndl
  write(u) g_z0hv
  write(u) g_z0qv
  write(u) g_forc_hgt_u_patch
  write(u) g_forc_hgt_t_patch
  write(u) g_forc_hgt_q_patch
  write(u) g_ram1
  write(u) g_rb1
  write(u) g_rah1
  write(u) g_rah2
  write(u) g_raw1
  write(u) g_raw2
  write(u) g_ustar
  write(u) g_um
  write(u) g_uaf
  write(u) g_taf
  write(u) g_qaf
  write(u) g_obu
  write(u) g_zeta
  write(u) g_vpd
  write(u) g_num_iter
  write(u) g_u10
  write(u) g_u10_clm
  write(u) g_va
  write(u) g_fv
  write(u) g_vds
  write(u) g_btran
  write(u) g_rresis
  write(u) g_rootr
  write(u) g_eflx_sh_veg
  write(u) g_eflx_sh_stem
  write(u) g_taux
  write(u) g_tauy
  write(u) g_ulrad
  write(u) g_dhsdt_canopy
  write(u) g_qflx_tran_veg
  write(u) g_qflx_evap_veg
  write(u) g_rh_af
  write(u) g_q_ref2m
  write(u) g_rh_ref2m
  write(u) g_rh_ref2m_r
  write(u) g_vpd_ref2m
  write(u) g_iwue_ln
  write(u) g_t_ref2m
  write(u) g_t_ref2m_r
  write(u) g_t_skin
  write(u) g_stem_biomass
  write(u) g_leaf_biomass
  write(u) g_dleaf_patch
  write(u) g_rssun
  write(u) g_rssha
  write(u) g_psnsun
  write(u) g_psnsha
  write(u) g_psnsun_wc
  write(u) g_psnsun_wj
  write(u) g_psnsun_wp
  write(u) g_psnsha_wc
  write(u) g_psnsha_wj
  write(u) g_psnsha_wp
  write(u) g_fpsn
  write(u) g_fpsn_wc
  write(u) g_fpsn_wj
  write(u) g_fpsn_wp
  write(u) g_lmrsun
  write(u) g_lmrsha
  write(u) g_gs_mol_sun
  write(u) g_gs_mol_sha
  write(u) g_gs_mol_sun_ln
  write(u) g_gs_mol_sha_ln
  write(u) g_cisun_z
  write(u) g_cisha_z
  write(u) g_rssun_z
  write(u) g_rssha_z
  write(u) g_lmrsun_z
  write(u) g_lmrsha_z
  write(u) g_psnsun_z
  write(u) g_psnsha_z
  write(u) g_vcmax_z
  write(u) g_tpu_z
  write(u) g_kp_z
  write(u) g_ac
  write(u) g_aj
  write(u) g_ap
  write(u) g_ag
  write(u) g_an
  write(u) g_gb_mol
  write(u) g_gs_mol
  write(u) g_cp
  write(u) g_kc
  write(u) g_ko
  write(u) g_qe
  write(u) g_bbb
  write(u) g_mbb
  write(u) g_rh_leaf
  write(u) g_vpd_can
  write(u) g_lnca
  write(u) g_c3flag
  write(u) g_canopy_err
  write(u) y_qflx_rain_plus_snomelt
  write(u) y_qflx_snow_h2osfc
  write(u) y_qflx_floodc
  write(u) y_qflx_liqevap_from_top_layer
  write(u) y_qflx_ev_soil
  write(u) y_qflx_ev_h2osfc
  write(u) y_zwt
  write(u) y_zwt_perched
  write(u) y_frost_table
  write(u) y_wtfact
  write(u) y_topo_slope
  write(u) y_h2osfc_thresh
  write(u) y_frac_h2osfc_nosnow
  write(u) y_fsat
  write(u) y_fcov
  write(u) y_qflx_sat_excess_surf
  write(u) y_qflx_top_soil
  write(u) y_qflx_in_soil
  write(u) y_qflx_top_soil_to_h2osfc
  write(u) y_qinmax
  write(u) y_qflx_infl_excess
  write(u) y_qflx_in_soil_limited
  write(u) y_qflx_in_h2osfc
  write(u) y_qflx_infl_excess_surf
  write(u) y_qflx_h2osfc_surf
  write(u) y_qflx_h2osfc_drain
  write(u) y_qflx_surf
  write(u) y_h2osoi_vol
  write(u) y_qflx_liqdew_to_top_layer
  write(u) y_qflx_soliddew_to_top_layer
  write(u) y_qflx_solidevap_from_top_layer
  write(u) y_qflx_snwcp_liq
  write(u) y_qflx_drain_perched
  write(u) y_qflx_latflow_out
  write(u) y_qflx_latflow_in
  write(u) y_qflx_rsub_sat
  write(u) y_qflx_ice_runoff_xs
  write(u) y_qflx_drain
  write(u) y_qflx_qrgwl
  write(u) y_qflx_runoff
  close(u)
end program drive_step
