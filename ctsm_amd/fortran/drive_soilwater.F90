! drive_soilwater.F90 -- SoilWater from Fortran through ctsm_gpu_mod.
!
! The column state is held the way CTSM holds it: one allocatable array per
! member with the reference's bounds (begc:endc, -nlevsno+1:nlevgrnd, ...),
! bound to the library by c_loc of its first element, with 1-based global
! column filters; ctsm_soilwater then runs the moisture-form Richards solve
! on those arrays in place (SoilWaterMovementMod.F90:240-243).
!
! Usage: drive_soilwater IN OUT.  IN holds ncol, nlevsno, nlevsoi, nlevgrnd
! (int32) and then every ctsm_col_state member in member order, each array in
! Fortran order; OUT receives the members after the call.  Prints the return
! code.
program drive_soilwater
  use iso_c_binding
  use ctsm_gpu_mod
  implicit none
  integer(c_int32_t) :: ncol, nlevsno, nlevsoi, nlevgrnd
  integer :: begc, endc, u, i, rc
  character(len=4096) :: fin, fout
  character(len=16) :: bcarg
  type(ctsm_config_t) :: cfg
  type(ctsm_layout_t) :: lay
  type(ctsm_col_state_t) :: col
  type(c_ptr) :: ctx
  integer(c_int32_t), allocatable :: filter(:)
  integer(c_int32_t), allocatable, target :: snl(:)
  integer(c_int32_t), allocatable, target :: nbedrock(:)
  integer(c_int32_t), allocatable, target :: patchi(:)
  integer(c_int32_t), allocatable, target :: patchf(:)
  real(c_double), allocatable, target :: dz(:,:)
  real(c_double), allocatable, target :: z(:,:)
  real(c_double), allocatable, target :: zi(:,:)
  real(c_double), allocatable, target :: watsat(:,:)
  real(c_double), allocatable, target :: bsw(:,:)
  real(c_double), allocatable, target :: sucsat(:,:)
  real(c_double), allocatable, target :: hksat(:,:)
  real(c_double), allocatable, target :: tkmg(:,:)
  real(c_double), allocatable, target :: tkdry(:,:)
  real(c_double), allocatable, target :: csol(:,:)
  real(c_double), allocatable, target :: eff_porosity(:,:)
  real(c_double), allocatable, target :: smp_l(:,:)
  real(c_double), allocatable, target :: hk_l(:,:)
  real(c_double), allocatable, target :: thk(:,:)
  real(c_double), allocatable, target :: icefrac(:,:)
  real(c_double), allocatable, target :: qcharge(:)
  real(c_double), allocatable, target :: num_substeps(:)
  real(c_double), allocatable, target :: qin(:,:)
  real(c_double), allocatable, target :: qout(:,:)
  real(c_double), allocatable, target :: h2osoi_liq(:,:)
  real(c_double), allocatable, target :: h2osoi_ice(:,:)
  real(c_double), allocatable, target :: h2osfc(:)
  real(c_double), allocatable, target :: h2osno_no_layers(:)
  real(c_double), allocatable, target :: int_snow(:)
  real(c_double), allocatable, target :: frac_sno_eff(:)
  real(c_double), allocatable, target :: frac_h2osfc(:)
  real(c_double), allocatable, target :: snow_depth(:)
  real(c_double), allocatable, target :: bw(:,:)
  real(c_double), allocatable, target :: snomelt_accum(:)
  real(c_double), allocatable, target :: qflx_infl(:)
  real(c_double), allocatable, target :: qflx_rootsoi(:,:)
  real(c_double), allocatable, target :: qflx_snomelt(:)
  real(c_double), allocatable, target :: qflx_snofrz(:)
  real(c_double), allocatable, target :: qflx_snow_drain(:)
  real(c_double), allocatable, target :: qflx_snomelt_lyr(:,:)
  real(c_double), allocatable, target :: qflx_snofrz_lyr(:,:)
  real(c_double), allocatable, target :: qflx_h2osfc_to_ice(:)
  real(c_double), allocatable, target :: t_soisno(:,:)
  real(c_double), allocatable, target :: t_h2osfc(:)
  real(c_double), allocatable, target :: t_grnd(:)
  real(c_double), allocatable, target :: fact(:,:)
  real(c_double), allocatable, target :: xmf(:)
  real(c_double), allocatable, target :: xmf_h2osfc(:)
  real(c_double), allocatable, target :: c_h2osfc(:)
  integer(c_int32_t), allocatable, target :: imelt(:,:)
  real(c_double), allocatable, target :: emg(:)
  real(c_double), allocatable, target :: htvp(:)
  real(c_double), allocatable, target :: eflx_bot(:)
  real(c_double), allocatable, target :: eflx_fgr12(:)
  real(c_double), allocatable, target :: eflx_fgr(:,:)
  real(c_double), allocatable, target :: eflx_snomelt(:)
  real(c_double), allocatable, target :: eflx_h2osfc_to_snow(:)
  real(c_double), allocatable, target :: forc_lwrad(:)
  real(c_double), allocatable, target :: zwt(:)
  real(c_double), allocatable, target :: smpmin(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, nlevsno, nlevsoi, nlevgrnd
  begc = 1
  endc = ncol
  allocate(snl(begc:endc))
  allocate(nbedrock(begc:endc))
  allocate(patchi(begc:endc))
  allocate(patchf(begc:endc))
  allocate(dz(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(z(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(zi(begc:endc, -nlevsno:nlevgrnd))
  allocate(watsat(begc:endc, 1:nlevgrnd))
  allocate(bsw(begc:endc, 1:nlevgrnd))
  allocate(sucsat(begc:endc, 1:nlevgrnd))
  allocate(hksat(begc:endc, 1:nlevgrnd))
  allocate(tkmg(begc:endc, 1:nlevgrnd))
  allocate(tkdry(begc:endc, 1:nlevgrnd))
  allocate(csol(begc:endc, 1:nlevgrnd))
  allocate(eff_porosity(begc:endc, 1:nlevgrnd))
  allocate(smp_l(begc:endc, 1:nlevgrnd))
  allocate(hk_l(begc:endc, 1:nlevgrnd))
  allocate(thk(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(icefrac(begc:endc, 1:nlevgrnd))
  allocate(qcharge(begc:endc))
  allocate(num_substeps(begc:endc))
  allocate(qin(begc:endc, 1:nlevgrnd))
  allocate(qout(begc:endc, 1:nlevgrnd))
  allocate(h2osoi_liq(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(h2osoi_ice(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(h2osfc(begc:endc))
  allocate(h2osno_no_layers(begc:endc))
  allocate(int_snow(begc:endc))
  allocate(frac_sno_eff(begc:endc))
  allocate(frac_h2osfc(begc:endc))
  allocate(snow_depth(begc:endc))
  allocate(bw(begc:endc, -nlevsno+1:0))
  allocate(snomelt_accum(begc:endc))
  allocate(qflx_infl(begc:endc))
  allocate(qflx_rootsoi(begc:endc, 1:nlevgrnd))
  allocate(qflx_snomelt(begc:endc))
  allocate(qflx_snofrz(begc:endc))
  allocate(qflx_snow_drain(begc:endc))
  allocate(qflx_snomelt_lyr(begc:endc, -nlevsno+1:0))
  allocate(qflx_snofrz_lyr(begc:endc, -nlevsno+1:0))
  allocate(qflx_h2osfc_to_ice(begc:endc))
  allocate(t_soisno(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(t_h2osfc(begc:endc))
  allocate(t_grnd(begc:endc))
  allocate(fact(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(xmf(begc:endc))
  allocate(xmf_h2osfc(begc:endc))
  allocate(c_h2osfc(begc:endc))
  allocate(imelt(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(emg(begc:endc))
  allocate(htvp(begc:endc))
  allocate(eflx_bot(begc:endc))
  allocate(eflx_fgr12(begc:endc))
  allocate(eflx_fgr(begc:endc, 1:nlevgrnd))
  allocate(eflx_snomelt(begc:endc))
  allocate(eflx_h2osfc_to_snow(begc:endc))
  allocate(forc_lwrad(begc:endc))
  allocate(zwt(begc:endc))
  allocate(smpmin(begc:endc))
  read(u) snl
  read(u) nbedrock
  read(u) patchi
  read(u) patchf
  read(u) dz
  read(u) z
  read(u) zi
  read(u) watsat
  read(u) bsw
  read(u) sucsat
  read(u) hksat
  read(u) tkmg
  read(u) tkdry
  read(u) csol
  read(u) eff_porosity
  read(u) smp_l
  read(u) hk_l
  read(u) thk
  read(u) icefrac
  read(u) qcharge
  read(u) num_substeps
  read(u) qin
  read(u) qout
  read(u) h2osoi_liq
  read(u) h2osoi_ice
  read(u) h2osfc
  read(u) h2osno_no_layers
  read(u) int_snow
  read(u) frac_sno_eff
  read(u) frac_h2osfc
  read(u) snow_depth
  read(u) bw
  read(u) snomelt_accum
  read(u) qflx_infl
  read(u) qflx_rootsoi
  read(u) qflx_snomelt
  read(u) qflx_snofrz
  read(u) qflx_snow_drain
  read(u) qflx_snomelt_lyr
  read(u) qflx_snofrz_lyr
  read(u) qflx_h2osfc_to_ice
  read(u) t_soisno
  read(u) t_h2osfc
  read(u) t_grnd
  read(u) fact
  read(u) xmf
  read(u) xmf_h2osfc
  read(u) c_h2osfc
  read(u) imelt
  read(u) emg
  read(u) htvp
  read(u) eflx_bot
  read(u) eflx_fgr12
  read(u) eflx_fgr
  read(u) eflx_snomelt
  read(u) eflx_h2osfc_to_snow
  read(u) forc_lwrad
  read(u) zwt
  read(u) smpmin
  close(u)

  col%snl = c_loc(snl(begc))
  col%nbedrock = c_loc(nbedrock(begc))
  col%patchi = c_loc(patchi(begc))
  col%patchf = c_loc(patchf(begc))
  col%dz = c_loc(dz(begc, -nlevsno+1))
  col%z = c_loc(z(begc, -nlevsno+1))
  col%zi = c_loc(zi(begc, -nlevsno))
  col%watsat = c_loc(watsat(begc, 1))
  col%bsw = c_loc(bsw(begc, 1))
  col%sucsat = c_loc(sucsat(begc, 1))
  col%hksat = c_loc(hksat(begc, 1))
  col%tkmg = c_loc(tkmg(begc, 1))
  col%tkdry = c_loc(tkdry(begc, 1))
  col%csol = c_loc(csol(begc, 1))
  col%eff_porosity = c_loc(eff_porosity(begc, 1))
  col%smp_l = c_loc(smp_l(begc, 1))
  col%hk_l = c_loc(hk_l(begc, 1))
  col%thk = c_loc(thk(begc, -nlevsno+1))
  col%icefrac = c_loc(icefrac(begc, 1))
  col%qcharge = c_loc(qcharge(begc))
  col%num_substeps = c_loc(num_substeps(begc))
  col%qin = c_loc(qin(begc, 1))
  col%qout = c_loc(qout(begc, 1))
  col%h2osoi_liq = c_loc(h2osoi_liq(begc, -nlevsno+1))
  col%h2osoi_ice = c_loc(h2osoi_ice(begc, -nlevsno+1))
  col%h2osfc = c_loc(h2osfc(begc))
  col%h2osno_no_layers = c_loc(h2osno_no_layers(begc))
  col%int_snow = c_loc(int_snow(begc))
  col%frac_sno_eff = c_loc(frac_sno_eff(begc))
  col%frac_h2osfc = c_loc(frac_h2osfc(begc))
  col%snow_depth = c_loc(snow_depth(begc))
  col%bw = c_loc(bw(begc, -nlevsno+1))
  col%snomelt_accum = c_loc(snomelt_accum(begc))
  col%qflx_infl = c_loc(qflx_infl(begc))
  col%qflx_rootsoi = c_loc(qflx_rootsoi(begc, 1))
  col%qflx_snomelt = c_loc(qflx_snomelt(begc))
  col%qflx_snofrz = c_loc(qflx_snofrz(begc))
  col%qflx_snow_drain = c_loc(qflx_snow_drain(begc))
  col%qflx_snomelt_lyr = c_loc(qflx_snomelt_lyr(begc, -nlevsno+1))
  col%qflx_snofrz_lyr = c_loc(qflx_snofrz_lyr(begc, -nlevsno+1))
  col%qflx_h2osfc_to_ice = c_loc(qflx_h2osfc_to_ice(begc))
  col%t_soisno = c_loc(t_soisno(begc, -nlevsno+1))
  col%t_h2osfc = c_loc(t_h2osfc(begc))
  col%t_grnd = c_loc(t_grnd(begc))
  col%fact = c_loc(fact(begc, -nlevsno+1))
  col%xmf = c_loc(xmf(begc))
  col%xmf_h2osfc = c_loc(xmf_h2osfc(begc))
  col%c_h2osfc = c_loc(c_h2osfc(begc))
  col%imelt = c_loc(imelt(begc, -nlevsno+1))
  col%emg = c_loc(emg(begc))
  col%htvp = c_loc(htvp(begc))
  col%eflx_bot = c_loc(eflx_bot(begc))
  col%eflx_fgr12 = c_loc(eflx_fgr12(begc))
  col%eflx_fgr = c_loc(eflx_fgr(begc, 1))
  col%eflx_snomelt = c_loc(eflx_snomelt(begc))
  col%eflx_h2osfc_to_snow = c_loc(eflx_h2osfc_to_snow(begc))
  col%forc_lwrad = c_loc(forc_lwrad(begc))
  col%zwt = c_loc(zwt(begc))
  col%smpmin = c_loc(smpmin(begc))

  call ctsm_config_default(cfg)
  cfg%nlevsno = nlevsno
  cfg%nlevsoi = nlevsoi
  cfg%nlevgrnd = nlevgrnd
  ! optional third argument: soilwater_movement's lower_boundary_condition
  if (command_argument_count() >= 3) then
     call get_command_argument(3, bcarg)
     read(bcarg, *) cfg%lower_boundary_condition
  end if
  rc = ctsm_create(cfg, 0, ctx)
  if (rc /= CTSM_OK) then
     print '(I0)', rc
     stop 1
  end if
  lay%begc_alloc = begc
  lay%ldc = endc - begc + 1
  lay%begp_alloc = 1
  lay%ldp = 0
  allocate(filter(ncol))
  filter = [(i, i = begc, endc)]
  rc = ctsm_soilwater(ctx, lay, col, ncol, filter, 1800.0_c_double)
  print '(I0)', rc
  i = ctsm_destroy(ctx)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) snl
  write(u) nbedrock
  write(u) patchi
  write(u) patchf
  write(u) dz
  write(u) z
  write(u) zi
  write(u) watsat
  write(u) bsw
  write(u) sucsat
  write(u) hksat
  write(u) tkmg
  write(u) tkdry
  write(u) csol
  write(u) eff_porosity
  write(u) smp_l
  write(u) hk_l
  write(u) thk
  write(u) icefrac
  write(u) qcharge
  write(u) num_substeps
  write(u) qin
  write(u) qout
  write(u) h2osoi_liq
  write(u) h2osoi_ice
  write(u) h2osfc
  write(u) h2osno_no_layers
  write(u) int_snow
  write(u) frac_sno_eff
  write(u) frac_h2osfc
  write(u) snow_depth
  write(u) bw
  write(u) snomelt_accum
  write(u) qflx_infl
  write(u) qflx_rootsoi
  write(u) qflx_snomelt
  write(u) qflx_snofrz
  write(u) qflx_snow_drain
  write(u) qflx_snomelt_lyr
  write(u) qflx_snofrz_lyr
  write(u) qflx_h2osfc_to_ice
  write(u) t_soisno
  write(u) t_h2osfc
  write(u) t_grnd
  write(u) fact
  write(u) xmf
  write(u) xmf_h2osfc
  write(u) c_h2osfc
  write(u) imelt
  write(u) emg
  write(u) htvp
  write(u) eflx_bot
  write(u) eflx_fgr12
  write(u) eflx_fgr
  write(u) eflx_snomelt
  write(u) eflx_h2osfc_to_snow
  write(u) forc_lwrad
  write(u) zwt
  write(u) smpmin
  close(u)
end program drive_soilwater
