! A second Fortran caller of the drop-in modules: derivatives (dfdx, dfdy, divuv),
! helmholtz (solve_nhom_helmholtz with the operators of the driver's mats module),
! sgs_models (sigma_model), filters (filter_q_3pt), dealias (setup_dealiasing,
! dealiasing_2d_fftw) and modhorfft (norm), on the device through lsbflow_amd, with the
! reference's use statements and argument lists.  Reads the grid sizes, sigma, two
! spectral fields u, v, a right-hand side, boundary data, a velocity gradient and two
! physical fields; writes every result.
program shim_check2
  use iso_c_binding
  use lsb_capi
  use lsb_context
  use modhorfft
  use derivatives
  use helmholtz
  use sgs_models
  use filters
  use dealias
  implicit none
  character(len=1024) :: inpath, outpath
  integer :: u, nx, ny, nz, nzp, nxpp, nzm
  real(c_double) :: sigma
  real(c_double), parameter :: pi = 3.141592653589793d0
  type(lsb_config) :: cfg
  type(c_ptr) :: ctx
  real(c_double), allocatable :: uf(:, :, :), vf(:, :, :), rhs(:, :, :), bc(:, :, :)
  real(c_double), allocatable :: dx(:, :, :), dy(:, :, :), dv(:, :, :)
  real(c_double), allocatable :: g(:, :, :, :, :), tau(:, :, :, :, :)
  real(c_double), allocatable :: pa(:, :, :), pb(:, :, :), pf(:, :, :), pab(:, :, :)
  real(c_double), allocatable :: n1(:, :, :), n2(:, :, :), n3(:, :, :), n4(:, :, :)
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
  allocate(uf(nx, ny, nzp), vf(nx, ny, nzp), rhs(nx, ny, nzp), bc(nx, ny, nzp))
  allocate(dx(nx, ny, nzp), dy(nx, ny, nzp), dv(nx, ny, nzp))
  allocate(g(nxpp, ny, nzp, 3, 3), tau(nxpp, ny, nzp, 3, 3))
  allocate(pa(nxpp, ny, nzp), pb(nxpp, ny, nzp), pf(nxpp, ny, nzp), pab(nxpp, ny, nzp))
  allocate(n1(nx, ny, nzp), n2(nx, ny, nzp), n3(nx, ny, nzp), n4(nx, ny, nzp))
  read(u) uf
  read(u) vf
  read(u) rhs
  read(u) bc
  read(u) g
  read(u) pa
  read(u) pb
  close(u)

  ! channel mapping 3, homogeneous walls, the sigma model, the padded-product dealiasing
  ! in the reference's ky placement, FP64 form
  cfg = lsb_config(nx, ny, nz, 3, 0, 0, 1, 2, 1, 1, 0, 0, 0, 0, 0, 0, LSB_FP64, 0, &
                   2.d0 * pi, pi, 2.d0, 0.d0, 1.d0, 1.d-3, 1.d0, 0.71d0, &
                   1.d0, 2.d0, 0.1d0, 0.6d0, 1.d0, 20.d0, 0.d0, 0.d0, 1.d-3)
  call lsb_check(lsb_create(cfg, ctx), 'lsb_create')
  call lsb_context_set(ctx)
  allocate(dor(nzp, nzp), d(nzp, nzp), d2(nzp, nzp), mdz(nzp, nzp), md2(nzp, nzp))
  allocate(diag(nzm), smat(nzm, nzm), simat(nzm, nzm))
  call lsb_check(lsb_get_mats(ctx, dor, d, d2, mdz, md2, diag, smat, simat), 'lsb_get_mats')

  call dfdx(uf, dx)
  call dfdy(uf, dy)
  call divuv(uf, vf, dv)
  call solve_nhom_helmholtz(rhs, bc, sigma, diag, smat, simat)
  call sigma_model(g, tau)
  call filter_q_3pt(pa, pf, 1, 1, 1.d0, 2.d0, 1)
  call setup_dealiasing(0)
  call dealiasing_2d_fftw(pa, pb, pab)
  n1 = uf
  n2 = vf
  n3 = dx
  n4 = dy
  call norm(n1, n2, n3, n4)

  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(u) nx_lg, ny_lg
  write(u) dx
  write(u) dy
  write(u) dv
  write(u) rhs
  write(u) tau
  write(u) pf
  write(u) pab
  write(u) n1
  write(u) n2
  write(u) n3
  write(u) n4
  close(u)
  call lsb_check(lsb_destroy(ctx), 'lsb_destroy')
end program shim_check2
