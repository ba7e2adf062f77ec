! drive_stress.F90 -- written by gen_drive_step.py from the member list of
! ctsm_amd/abi.py; do not edit by hand.
!
! Plant-hydraulic-stress entries from Fortran through ctsm_gpu_mod, on
! arrays held as CTSM holds them (reference bounds, bound by c_loc).
!
! Usage: drive_stress IN OUT.  IN as drive_phs's (mode ignored), the tables
! followed by kmax (npft, 4), then every ctsm_phs_stress_state member; calls
! ctsm_phs_calcstress over every patch.
program drive_stress
  use iso_c_binding
  use ctsm_gpu_mod
  implicit none
  integer(c_int32_t) :: ncol, npatch, nlevsno, nlevsoi, nlevgrnd, npft, mode
  integer :: begc, endc, begp, endp, u, i, rc
  character(len=4096) :: fin, fout
  type(ctsm_config_t) :: cfg
  type(ctsm_layout_t) :: lay
  type(ctsm_phs_stress_state_t) :: phs
  type(ctsm_phs_params_t) :: prm
  type(c_ptr) :: ctx
  integer(c_int32_t), allocatable :: filter(:)
  real(c_double), allocatable, target :: krmax(:), root_radius(:), root_density(:), froot_leaf(:)
  real(c_double), allocatable, target :: psi50(:,:), ck(:,:)
  real(c_double), allocatable, target :: kmax(:,:)
  integer(c_int32_t), allocatable, target :: a_itype(:)
  integer(c_int32_t), allocatable, target :: a_column(:)
  integer(c_int32_t), allocatable, target :: a_local_morning(:)
  real(c_double), allocatable, target :: a_laisun(:)
  real(c_double), allocatable, target :: a_laisha(:)
  real(c_double), allocatable, target :: a_elai(:)
  real(c_double), allocatable, target :: a_esai(:)
  real(c_double), allocatable, target :: a_tsai(:)
  real(c_double), allocatable, target :: a_htop(:)
  real(c_double), allocatable, target :: a_fdry(:)
  real(c_double), allocatable, target :: a_thm(:)
  real(c_double), allocatable, target :: a_gb_mol(:)
  real(c_double), allocatable, target :: a_gs_mol_sun(:)
  real(c_double), allocatable, target :: a_gs_mol_sha(:)
  real(c_double), allocatable, target :: a_qsatl(:)
  real(c_double), allocatable, target :: a_qaf(:)
  real(c_double), allocatable, target :: a_k_soil_root(:,:)
  real(c_double), allocatable, target :: a_vegwp(:,:)
  real(c_double), allocatable, target :: a_vegwp_pd(:,:)
  real(c_double), allocatable, target :: a_bsun(:)
  real(c_double), allocatable, target :: a_bsha(:)
  real(c_double), allocatable, target :: a_qflx_tran_veg(:)
  real(c_double), allocatable, target :: a_forc_rho(:)
  real(c_double), allocatable, target :: a_forc_pbot(:)
  real(c_double), allocatable, target :: a_z(:,:)
  real(c_double), allocatable, target :: a_smp_l(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, npatch, nlevsno, nlevsoi, nlevgrnd, npft, mode
  allocate(krmax(0:npft-1), root_radius(0:npft-1), root_density(0:npft-1), froot_leaf(0:npft-1))
  allocate(psi50(0:npft-1, 4), ck(0:npft-1, 4))
  read(u) krmax, root_radius, root_density, froot_leaf, psi50, ck
  allocate(kmax(0:npft-1, 4))
  read(u) kmax
  prm%kmax = c_loc(kmax(0, 1))
  begc = 1
  endc = ncol
  begp = 1
  endp = npatch
  allocate(a_itype(begp:endp))
  allocate(a_column(begp:endp))
  allocate(a_local_morning(begp:endp))
  allocate(a_laisun(begp:endp))
  allocate(a_laisha(begp:endp))
  allocate(a_elai(begp:endp))
  allocate(a_esai(begp:endp))
  allocate(a_tsai(begp:endp))
  allocate(a_htop(begp:endp))
  allocate(a_fdry(begp:endp))
  allocate(a_thm(begp:endp))
  allocate(a_gb_mol(begp:endp))
  allocate(a_gs_mol_sun(begp:endp))
  allocate(a_gs_mol_sha(begp:endp))
  allocate(a_qsatl(begp:endp))
  allocate(a_qaf(begp:endp))
  allocate(a_k_soil_root(begp:endp, 1:nlevgrnd))
  allocate(a_vegwp(begp:endp, 1:4))
  allocate(a_vegwp_pd(begp:endp, 1:4))
  allocate(a_bsun(begp:endp))
  allocate(a_bsha(begp:endp))
  allocate(a_qflx_tran_veg(begp:endp))
  allocate(a_forc_rho(begc:endc))
  allocate(a_forc_pbot(begc:endc))
  allocate(a_z(begc:endc, -nlevsno+1:nlevgrnd))
  allocate(a_smp_l(begc:endc, 1:nlevgrnd))
  read(u) a_itype
  read(u) a_column
  read(u) a_local_morning
  read(u) a_laisun
  read(u) a_laisha
  read(u) a_elai
  read(u) a_esai
  read(u) a_tsai
  read(u) a_htop
  read(u) a_fdry
  read(u) a_thm
  read(u) a_gb_mol
  read(u) a_gs_mol_sun
  read(u) a_gs_mol_sha
  read(u) a_qsatl
  read(u) a_qaf
  read(u) a_k_soil_root
  read(u) a_vegwp
  read(u) a_vegwp_pd
  read(u) a_bsun
  read(u) a_bsha
  read(u) a_qflx_tran_veg
  read(u) a_forc_rho
  read(u) a_forc_pbot
  read(u) a_z
  read(u) a_smp_l
  close(u)
  phs%itype = c_loc(a_itype(begp))
  phs%column = c_loc(a_column(begp))
  phs%local_morning = c_loc(a_local_morning(begp))
  phs%laisun = c_loc(a_laisun(begp))
  phs%laisha = c_loc(a_laisha(begp))
  phs%elai = c_loc(a_elai(begp))
  phs%esai = c_loc(a_esai(begp))
  phs%tsai = c_loc(a_tsai(begp))
  phs%htop = c_loc(a_htop(begp))
  phs%fdry = c_loc(a_fdry(begp))
  phs%thm = c_loc(a_thm(begp))
  phs%gb_mol = c_loc(a_gb_mol(begp))
  phs%gs_mol_sun = c_loc(a_gs_mol_sun(begp))
  phs%gs_mol_sha = c_loc(a_gs_mol_sha(begp))
  phs%qsatl = c_loc(a_qsatl(begp))
  phs%qaf = c_loc(a_qaf(begp))
  phs%k_soil_root = c_loc(a_k_soil_root(begp, 1))
  phs%vegwp = c_loc(a_vegwp(begp, 1))
  phs%vegwp_pd = c_loc(a_vegwp_pd(begp, 1))
  phs%bsun = c_loc(a_bsun(begp))
  phs%bsha = c_loc(a_bsha(begp))
  phs%qflx_tran_veg = c_loc(a_qflx_tran_veg(begp))
  phs%forc_rho = c_loc(a_forc_rho(begc))
  phs%forc_pbot = c_loc(a_forc_pbot(begc))
  phs%z = c_loc(a_z(begc, -nlevsno+1))
  phs%smp_l = c_loc(a_smp_l(begc, 1))
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
  if (rc == CTSM_OK) then
     allocate(filter(npatch))
     filter = [(i, i = begp, endp)]
     rc = ctsm_phs_calcstress(ctx, lay, phs, npatch, filter)
  end if
  print '(I0)', rc
  i = ctsm_destroy(ctx)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) a_itype
  write(u) a_column
  write(u) a_local_morning
  write(u) a_laisun
  write(u) a_laisha
  write(u) a_elai
  write(u) a_esai
  write(u) a_tsai
  write(u) a_htop
  write(u) a_fdry
  write(u) a_thm
  write(u) a_gb_mol
  write(u) a_gs_mol_sun
  write(u) a_gs_mol_sha
  write(u) a_qsatl
  write(u) a_qaf
  write(u) a_k_soil_root
  write(u) a_vegwp
  write(u) a_vegwp_pd
  write(u) a_bsun
  write(u) a_bsha
  write(u) a_qflx_tran_veg
  write(u) a_forc_rho
  write(u) a_forc_pbot
  write(u) a_z
  write(u) a_smp_l
  close(u)
end program drive_stress
