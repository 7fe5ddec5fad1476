! Calls liblsbflow_amd.so through the lsb_capi interface module the way a
! Fortran driver would: the host-side setup routines (chebyshev.f90:9-48,
! chebyshev.f90:204-270, helmholtz.f90:13-126), the Orr-Sommerfeld solve with
! complex arguments (orrsomm.f90:110-562) and the error path of lsb_create.
! Reads its inputs from and writes its outputs to stream files.
program binding_check
  use iso_c_binding
  use lsb_capi
  implicit none
  integer(c_int) :: nzp, nzm, n1, n2, st(4), nzq, sto
  real(c_double) :: xl, xnu, eps, alphax
  real(c_double), allocatable :: dq(:,:), d2q(:,:), ubar(:), ubarp(:), ubarpp(:), zq(:)
  complex(c_double_complex) :: omega
  complex(c_double_complex), allocatable :: uvec(:), vvec(:), wvec(:)
  real(c_double), allocatable :: gf(:), y(:,:,:), dydx(:,:,:)
  real(c_double), allocatable :: dor(:,:), d(:,:), d2(:,:), dint(:,:), lam(:), e(:,:), ei(:,:)
  type(lsb_config) :: cfg
  type(c_ptr) :: ctx
  character(len=1024) :: inpath, outpath
  integer :: u

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  open(newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(u) nzp, n1, n2
  read(u) xl
  nzm = nzp - 2
  allocate(gf(nzp), y(n1, n2, nzp), dydx(n1, n2, nzp))
  allocate(dor(nzp, nzp), d(nzp, nzp), d2(nzp, nzp), dint(nzm, nzm), lam(nzm), e(nzm, nzm), ei(nzm, nzm))
  read(u) gf
  read(u) y
  read(u) nzq
  allocate(dq(nzq, nzq), d2q(nzq, nzq), ubar(nzq), ubarp(nzq), ubarpp(nzq), zq(nzq))
  allocate(uvec(nzq), vvec(nzq), wvec(nzq))
  read(u) dq, d2q, ubar, ubarp, ubarpp, zq
  read(u) xnu, eps, alphax, omega
  close(u)

  st(1) = lsb_build_cheb_zder_matrix(nzp, xl, gf, dor, d, d2)
  st(2) = lsb_dct_diff(n1, n2, nzp, xl, gf, y, dydx, 1_c_int)
  dint = d2(2:nzp-1, 2:nzp-1)
  st(3) = lsb_diagonalize_matrix(dint, nzm, lam, e, ei)
  sto = lsb_solve_orrsomm(nzq, dq, d2q, ubar, ubarp, ubarpp, zq, xnu, eps, alphax, 0.d0, omega, 0_c_int, &
                          uvec, vvec, wvec)

  ! an invalid configuration (odd nx) is refused before any device call
  cfg = lsb_config(3, 4, 8, 3, 0, 0, 0, 0, 2, 0, 0, 0, 0, 0, 1, 0, LSB_FP64, 0, &
                   6.283185307179586d0, 3.141592653589793d0, 2.d0, 0.d0, 1.d0, 1.d-3, 1.d0, 1.d0, &
                   1.d0, 2.d0, 0.1d0, 0.6d0, 1.d0, 20.d0, 0.d0, 0.d0, 1.d-3)
  st(4) = lsb_create(cfg, ctx)

  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(u) st, int(c_sizeof(cfg), c_int)
  write(u) dor, d, d2, dydx, lam, e, ei
  write(u) sto, omega, uvec, vvec, wvec
  close(u)
  write(*, '(a)') lsb_error_message()
end program binding_check
