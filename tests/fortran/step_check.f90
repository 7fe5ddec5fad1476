! A Fortran driver of the device-resident step (tier B), as the reference's time loop
! would call it: lsb_create with a configuration read as the C struct, the base flow
! (paral) installed, the state loaded once, nsteps steps enqueued, the state read back.
! Reads the configuration, nsteps, the base profiles and the initial u, v, w; writes the
! step counter, the time and the nine state fields.
program step_check
  use iso_c_binding
  use lsb_capi
  implicit none
  character(len=1024) :: inpath, outpath
  integer :: u, nsteps, nzp, nspec
  integer(c_int) :: istep
  real(c_double) :: t
  type(lsb_config) :: cfg
  type(c_ptr) :: ctx
  real(c_double), allocatable :: ubar(:), ubarp(:), ubarpp(:), zero1(:)
  real(c_double), allocatable :: uf(:), vf(:), wf(:), pf(:), pnf(:), unif(:), unmif(:), nlnif(:), nlnmif(:)

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  open(newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(u) cfg
  read(u) nsteps
  nzp = cfg%nz + 1
  nspec = cfg%nx * cfg%ny * nzp
  allocate(ubar(nzp), ubarp(nzp), ubarpp(nzp), zero1(nzp))
  allocate(uf(nspec), vf(nspec), wf(nspec), pf(nspec), pnf(nspec))
  allocate(unif(3 * nspec), unmif(3 * nspec), nlnif(3 * nspec), nlnmif(3 * nspec))
  read(u) ubar
  read(u) ubarp
  read(u) ubarpp
  read(u) uf
  read(u) vf
  read(u) wf
  close(u)
  zero1 = 0.d0
  pf = 0.d0
  pnf = 0.d0
  unif = 0.d0
  unmif = 0.d0
  nlnif = 0.d0
  nlnmif = 0.d0

  call lsb_check(lsb_create(cfg, ctx), 'lsb_create')
  call lsb_check(lsb_set_paral(ctx, ubar, ubarp, ubarpp, zero1, zero1), 'lsb_set_paral')
  call lsb_check(lsb_state_set(ctx, uf, vf, wf, pf, pnf, unif, unmif, nlnif, nlnmif, 0_c_int, 0.d0), &
                 'lsb_state_set')
  call lsb_check(lsb_timestep(ctx, int(nsteps, c_int)), 'lsb_timestep')
  call lsb_check(lsb_state_get(ctx, uf, vf, wf, pf, pnf, unif, unmif, nlnif, nlnmif, istep, t), &
                 'lsb_state_get')

  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(u) istep
  write(u) t
  write(u) uf
  write(u) vf
  write(u) wf
  write(u) pf
  write(u) pnf
  write(u) unif
  write(u) unmif
  write(u) nlnif
  write(u) nlnmif
  close(u)
  call lsb_check(lsb_destroy(ctx), 'lsb_destroy')
end program step_check
