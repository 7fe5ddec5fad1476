! A Fortran caller of the drop-in modules: the same use statements and calls as the
! reference's driver (horfft, dfdz, solve_helmholtz with the operator arguments of its
! mats module), on the device through lsbflow_amd.  Reads the grid sizes, sigma, one
! physical field and one spectral right-hand side; writes the transformed field, its
! z derivative, the solved right-hand side and the inverse transform back to physical.
program shim_check
  use iso_c_binding
  use lsb_capi
  use lsb_context
  use modhorfft
  use derivatives
  use helmholtz
  implicit none
  character(len=1024) :: inpath, outpath
  integer :: u, nx, ny, nz, nzp, nxpp, nzm
  real(c_double) :: sigma
  real(c_double), parameter :: pi = 3.141592653589793d0
  type(lsb_config) :: cfg
  type(c_ptr) :: ctx
  real(c_double), allocatable :: ph(:, :, :), back(:, :, :), sp(:, :, :), dz(:, :, :), rhs(:, :, :)
  real(c_double), allocatable :: dor(:, :), d(:, :), d2(:, :), mdz(:, :), md2(:, :)
  real(c_double), allocatable :: diag(:), smat(:, :), simat(:, :)

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  open(newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(u) nx, ny, nz
  read(u) sigma
  nzp = nz + 1
  nxpp = nx + 2
  nzm = nz - 1
  allocate(ph(nxpp, ny, nzp), back(nxpp, ny, nzp), sp(nx, ny, nzp), dz(nx, ny, nzp), rhs(nx, ny, nzp))
  read(u) ph
  read(u) rhs
  close(u)

  ! channel mapping 3, homogeneous walls, spectral filter dealiasing, FP64 form
  cfg = lsb_config(nx, ny, nz, 3, 0, 0, 0, 0, 2, 1, 0, 0, 0, 0, 0, 0, LSB_FP64, 0, &
                   2.d0 * pi, pi, 2.d0, 0.d0, 1.d0, 1.d-3, 1.d0, 0.71d0, &
                   1.d0, 2.d0, 0.1d0, 0.6d0, 1.d0, 20.d0, 0.d0, 0.d0, 1.d-3)
  call lsb_check(lsb_create(cfg, ctx), 'lsb_create')
  call lsb_context_set(ctx)

  ! the driver's mats module: the operators the context was built with
  allocate(dor(nzp, nzp), d(nzp, nzp), d2(nzp, nzp), mdz(nzp, nzp), md2(nzp, nzp))
  allocate(diag(nzm), smat(nzm, nzm), simat(nzm, nzm))
  call lsb_check(lsb_get_mats(ctx, dor, d, d2, mdz, md2, diag, smat, simat), 'lsb_get_mats')

  call horfft(ph, sp, -1)
  call dfdz(sp, dz)
  call solve_helmholtz(rhs, sigma, nzm, diag, smat, simat)
  back = 0.d0
  call horfft(back, sp, 1)

  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(u) sp
  write(u) dz
  write(u) rhs
  write(u) back
  close(u)
  call lsb_check(lsb_destroy(ctx), 'lsb_destroy')
end program shim_check
