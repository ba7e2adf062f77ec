! drive_phs.F90 -- written by gen_drive_step.py from the member list of
! ctsm_amd/abi.py; do not edit by hand.
!
! Plant-hydraulic-stress entries from Fortran through ctsm_gpu_mod, on
! arrays held as CTSM holds them (reference bounds, bound by c_loc).
!
! Usage: drive_phs IN OUT.  IN holds, as int32, ncol, npatch, nlevsno, nlevsoi,
! nlevgrnd, npft and mode (1 conductances over every patch, 2 the sink over
! every column, 3 both); the PHS tables (krmax, root_radius, root_density,
! froot_leaf: npft doubles each; psi50 and ck: (npft, 4)); then every
! ctsm_phs_state member in member order.  OUT receives the members after the
! calls.  Prints the return code.
program drive_phs
  use iso_c_binding
  use ctsm_gpu_mod
  implicit none
  integer(c_int32_t) :: ncol, npatch, nlevsno, nlevsoi, nlevgrnd, npft, mode
  integer :: begc, endc, begp, endp, u, i, rc
  character(len=4096) :: fin, fout
  type(ctsm_config_t) :: cfg
  type(ctsm_layout_t) :: lay
  type(ctsm_phs_state_t) :: phs
  type(ctsm_phs_params_t) :: prm
  type(c_ptr) :: ctx
  integer(c_int32_t), allocatable :: filter(:)
  real(c_double), allocatable, target :: krmax(:), root_radius(:), root_density(:), froot_leaf(:)
  real(c_double), allocatable, target :: psi50(:,:), ck(:,:)

  integer(c_int32_t), allocatable, target :: a_itype(:)
  integer(c_int32_t), allocatable, target :: a_column(:)
  integer(c_int32_t), allocatable, target :: a_active(:)
  integer(c_int32_t), allocatable, target :: a_frac_veg_nosno(:)
  real(c_double), allocatable, target :: a_wtcol(:)
  real(c_double), allocatable, target :: a_froot_carbon(:)
  real(c_double), allocatable, target :: a_tsai(:)
  real(c_double), allocatable, target :: a_tlai(:)
  real(c_double), allocatable, target :: a_rootfr(:,:)
  real(c_double), allocatable, target :: a_vegwp(:,:)
  real(c_double), allocatable, target :: a_k_soil_root(:,:)
  real(c_double), allocatable, target :: a_root_conductance(:,:)
  real(c_double), allocatable, target :: a_soil_conductance(:,:)
  real(c_double), allocatable, target :: a_qflx_hydr_redist(:)
  integer(c_int32_t), allocatable, target :: a_patchi(:)
  integer(c_int32_t), allocatable, target :: a_patchf(:)
  real(c_double), allocatable, target :: a_dz(:,:)
  real(c_double), allocatable, target :: a_z(:,:)
  real(c_double), allocatable, target :: a_hksat(:,:)
  real(c_double), allocatable, target :: a_hk_l(:,:)
  real(c_double), allocatable, target :: a_smp_l(:,:)
  real(c_double), allocatable, target :: a_qflx_rootsoi(:,:)
  real(c_double), allocatable, target :: a_qflx_phs_neg(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, npatch, nlevsno, nlevsoi, nlevgrnd, npft, mode
  allocate(krmax(0:npft-1), root_radius(0:npft-1), root_density(0:npft-1), froot_leaf(0:npft-1))
  allocate(psi50(0:npft-1, 4), ck(0:npft-1, 4))
  read(u) krmax, root_radius, root_density, froot_leaf, psi50, ck

  begc = 1
  endc = ncol
  begp = 1
  endp = npatch
  allocate(a_itype(begp:endp))
  allocate(a_column(begp:endp))
  allocate(a_active(begp:endp))
  allocate(a_frac_veg_nosno(begp:endp))
  allocate(a_wtcol(begp:endp))
  allocate(a_froot_carbon(begp:endp))
  allocate(a_tsai(begp:endp))
  allocate(a_tlai(begp:endp))
  allocate(a_rootfr(begp:endp, 1:nlevgrnd))
  allocate(a_vegwp(begp:endp, 1:4))
  allocate(a_k_soil_root(begp:endp, 1:nlevgrnd))
  allocate(a_root_conductance(begp:endp, 1:nlevgrnd))
  allocate(a_soil_conductance(begp:endp, 1:nlevgrnd))
  allocate(a_qflx_hydr_redist(begp:endp))
  allocate(a_patchi(begc:endc))
  allocate(a_patchf(begc:endc))
  allocate(a_dz(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(a_z(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(a_hksat(begc:endc, 1:nlevgrnd))
  allocate(a_hk_l(begc:endc, 1:nlevgrnd))
  allocate(a_smp_l(begc:endc, 1:nlevgrnd))
  allocate(a_qflx_rootsoi(begc:endc, 1:nlevgrnd))
  allocate(a_qflx_phs_neg(begc:endc))
  read(u) a_itype
  read(u) a_column
  read(u) a_active
  read(u) a_frac_veg_nosno
  read(u) a_wtcol
  read(u) a_froot_carbon
  read(u) a_tsai
  read(u) a_tlai
  read(u) a_rootfr
  read(u) a_vegwp
  read(u) a_k_soil_root
  read(u) a_root_conductance
  read(u) a_soil_conductance
  read(u) a_qflx_hydr_redist
  read(u) a_patchi
  read(u) a_patchf
  read(u) a_dz
  read(u) a_z
  read(u) a_hksat
  read(u) a_hk_l
  read(u) a_smp_l
  read(u) a_qflx_rootsoi
  read(u) a_qflx_phs_neg
  close(u)
  phs%itype = c_loc(a_itype(begp))
  phs%column = c_loc(a_column(begp))
  phs%active = c_loc(a_active(begp))
  phs%frac_veg_nosno = c_loc(a_frac_veg_nosno(begp))
  phs%wtcol = c_loc(a_wtcol(begp))
  phs%froot_carbon = c_loc(a_froot_carbon(begp))
  phs%tsai = c_loc(a_tsai(begp))
  phs%tlai = c_loc(a_tlai(begp))
  phs%rootfr = c_loc(a_rootfr(begp, 1))
  phs%vegwp = c_loc(a_vegwp(begp, 1))
  phs%k_soil_root = c_loc(a_k_soil_root(begp, 1))
  phs%root_conductance = c_loc(a_root_conductance(begp, 1))
  phs%soil_conductance = c_loc(a_soil_conductance(begp, 1))
  phs%qflx_hydr_redist = c_loc(a_qflx_hydr_redist(begp))
  phs%patchi = c_loc(a_patchi(begc))
  phs%patchf = c_loc(a_patchf(begc))
  phs%dz = c_loc(a_dz(begc, -nlevsno+1))
  phs%z = c_loc(a_z(begc, -nlevsno+1))
  phs%hksat = c_loc(a_hksat(begc, 1))
  phs%hk_l = c_loc(a_hk_l(begc, 1))
  phs%smp_l = c_loc(a_smp_l(begc, 1))
  phs%qflx_rootsoi = c_loc(a_qflx_rootsoi(begc, 1))
  phs%qflx_phs_neg = c_loc(a_qflx_phs_neg(begc))
  prm%npft = npft
  prm%krmax = c_loc(krmax(0))
  prm%root_radius = c_loc(root_radius(0))
  prm%root_density = c_loc(root_density(0))
  prm%froot_leaf = c_loc(froot_leaf(0))
  prm%psi50 = c_loc(psi50(0, 1))
  prm%ck = c_loc(ck(0, 1))

  call ctsm_config_default(cfg)
  cfg%nlevsno = nlevsno
  cfg%nlevsoi = nlevsoi
  cfg%nlevgrnd = nlevgrnd
  rc = ctsm_create(cfg, 0, ctx)
  if (rc /= CTSM_OK) then
     print '(I0)', rc
     stop 1
  end if
  lay%begc_alloc = begc
  lay%ldc = endc - begc + 1
  lay%begp_alloc = begp
  lay%ldp = endp - begp + 1
  lay%begg_alloc = 1
  lay%ldg = 1
  rc = ctsm_set_phs_params(ctx, prm)
  if (rc == CTSM_OK .and. iand(mode, 1) /= 0) then
     allocate(filter(npatch))
     filter = [(i, i = begp, endp)]
     rc = ctsm_phs_soil_root_conductance(ctx, lay, phs, npatch, filter)
     deallocate(filter)
  end if
  if (rc == CTSM_OK .and. iand(mode, 2) /= 0) then
     allocate(filter(ncol))
     filter = [(i, i = begc, endc)]
     rc = ctsm_phs_root_sink(ctx, lay, phs, ncol, filter)
  end if
  print '(I0)', rc
  i = ctsm_destroy(ctx)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) a_itype
  write(u) a_column
  write(u) a_active
  write(u) a_frac_veg_nosno
  write(u) a_wtcol
  write(u) a_froot_carbon
  write(u) a_tsai
  write(u) a_tlai
  write(u) a_rootfr
  write(u) a_vegwp
  write(u) a_k_soil_root
  write(u) a_root_conductance
  write(u) a_soil_conductance
  write(u) a_qflx_hydr_redist
  write(u) a_patchi
  write(u) a_patchf
  write(u) a_dz
  write(u) a_z
  write(u) a_hksat
  write(u) a_hk_l
  write(u) a_smp_l
  write(u) a_qflx_rootsoi
  write(u) a_qflx_phs_neg
  close(u)
end program drive_phs
