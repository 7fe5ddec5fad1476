! Drop-in modules with the reference's module and procedure names: modhorfft
! (horfft.f90), derivatives (derivatives.f90), helmholtz (helmholtz.f90), sgs_models
! (sgsterm.f90), filters (filters.f90) and dealias (dealiasing.f90).  A driver compiled
! against these modules instead of the reference's keeps its call sites unchanged and
! runs each call on the GPU through the C-ABI (lsb_capi); the reference is compiled
! with 8-byte default reals (-fdefault-real-8), the shims take real(c_double) arrays.
!
! Arrays keep the reference's shapes and meaning and are passed as assumed-size
! arrays: their sizes are those of the context (the driver's dim module values given
! to lsb_create).  The context is created once by the driver after its own set-up and
! handed to lsb_context_set; every shim stops with the library's message when a call
! fails (the reference prints and continues or calls MPI_FINALIZE).  The operator
! arguments of the Helmholtz solves (diag, smat, simat) are passed through: the
! library uses them when they differ from the context's (lsb_set_mats).

module lsb_context
  use iso_c_binding
  use lsb_capi
  implicit none
  type(c_ptr), save :: g_ctx = c_null_ptr
contains
  subroutine lsb_context_set(ctx)
    type(c_ptr), intent(in) :: ctx
    g_ctx = ctx
  end subroutine lsb_context_set

  function lsb_ctx() result(ctx)
    type(c_ptr) :: ctx
    if (.not. c_associated(g_ctx)) then
      write(*, *) 'lsbflow_amd: no context (call lsb_context_set after lsb_create)'
      error stop 1
    end if
    ctx = g_ctx
  end function lsb_ctx
end module lsb_context

! horfft.f90: 2-D transforms of all planes and the spectral normalisation
module modhorfft
  use iso_c_binding
  use lsb_capi
  use lsb_context
  implicit none
  private
  public :: horfft, norm, dataflip_xtoy, dataflip_ytox, test_fft
contains
  ! horfft.f90:61-150; is = -1 physical -> spectral, +1 spectral -> physical
  subroutine horfft(ain, ainf, is)
    real(c_double), intent(inout) :: ain(*), ainf(*)
    integer, intent(in) :: is
    call lsb_check(lsb_horfft(lsb_ctx(), ain, ainf, int(is, c_int), 1_c_int), 'horfft')
  end subroutine horfft

  ! horfft.f90:153-224
  subroutine norm(uf, vf, wf, tempf)
    real(c_double), intent(inout) :: uf(*), vf(*), wf(*), tempf(*)
    call lsb_check(lsb_norm(lsb_ctx(), uf, vf, wf, tempf), 'norm')
  end subroutine norm

  ! horfft.f90:261-356 (one process)
  subroutine dataflip_xtoy(x, xf)
    real(c_double), intent(in) :: x(*)
    real(c_double), intent(out) :: xf(*)
    call lsb_check(lsb_dataflip_xtoy(lsb_ctx(), x, xf), 'dataflip_xtoy')
  end subroutine dataflip_xtoy

  ! horfft.f90:359-454 (one process)
  subroutine dataflip_ytox(x, xf)
    real(c_double), intent(inout) :: x(*)
    real(c_double), intent(in) :: xf(*)
    call lsb_check(lsb_dataflip_ytox(lsb_ctx(), x, xf), 'dataflip_ytox')
  end subroutine dataflip_ytox

  ! horfft.f90:7-58
  subroutine test_fft(u, avgerror)
    real(c_double), intent(in) :: u(*)
    real(c_double), intent(out) :: avgerror
    call lsb_check(lsb_test_fft(lsb_ctx(), u, avgerror), 'test_fft')
  end subroutine test_fft
end module modhorfft

! derivatives.f90: horizontal and wall-normal derivatives
module derivatives
  use iso_c_binding
  use lsb_capi
  use lsb_context
  implicit none
  private
  public :: zinteg, dfdz, dpdz, d2fdz2, dfdx, dfdy, lapl, div, divcheck, divuv
contains
  ! derivatives.f90:9-70
  subroutine zinteg(tmp, tint, n1, n2, mapping, per, rev)
    integer, intent(in) :: n1, n2, mapping, per, rev
    real(c_double), intent(in) :: tmp(*)
    real(c_double), intent(out) :: tint(*)
    call lsb_check(lsb_zinteg(lsb_ctx(), tmp, tint, int(n1, c_int), int(n2, c_int), int(mapping, c_int), &
                              int(per, c_int), int(rev, c_int)), 'zinteg')
  end subroutine zinteg

  ! derivatives.f90:112-132
  subroutine dfdz(uf, dudzf)
    real(c_double), intent(in) :: uf(*)
    real(c_double), intent(out) :: dudzf(*)
    call lsb_check(lsb_dfdz(lsb_ctx(), uf, dudzf), 'dfdz')
  end subroutine dfdz

  ! derivatives.f90:135-158; n is the context's nzp
  subroutine dpdz(n1, n2, n, u, dudz, p)
    integer, intent(in) :: n1, n2, n, p
    real(c_double), intent(in) :: u(*)
    real(c_double), intent(out) :: dudz(*)
    call lsb_check(lsb_dpdz(lsb_ctx(), int(n1, c_int), int(n2, c_int), u, dudz, int(p, c_int)), 'dpdz')
  end subroutine dpdz

  ! derivatives.f90:161-188
  subroutine d2fdz2(uf, dudz2f)
    real(c_double), intent(in) :: uf(*)
    real(c_double), intent(out) :: dudz2f(*)
    call lsb_check(lsb_d2fdz2(lsb_ctx(), uf, dudz2f), 'd2fdz2')
  end subroutine d2fdz2

  ! derivatives.f90:261-280
  subroutine dfdx(uf, dudxf)
    real(c_double), intent(in) :: uf(*)
    real(c_double), intent(out) :: dudxf(*)
    call lsb_check(lsb_dfdx(lsb_ctx(), uf, dudxf), 'dfdx')
  end subroutine dfdx

  ! derivatives.f90:283-303
  subroutine dfdy(uf, dudyf)
    real(c_double), intent(in) :: uf(*)
    real(c_double), intent(out) :: dudyf(*)
    call lsb_check(lsb_dfdy(lsb_ctx(), uf, dudyf), 'dfdy')
  end subroutine dfdy

  ! derivatives.f90:306-326
  subroutine lapl(uf, lapluf)
    real(c_double), intent(in) :: uf(*)
    real(c_double), intent(out) :: lapluf(*)
    call lsb_check(lsb_lapl(lsb_ctx(), uf, lapluf), 'lapl')
  end subroutine lapl

  ! derivatives.f90:329-346
  subroutine div(uf, vf, wf, divf)
    real(c_double), intent(in) :: uf(*), vf(*), wf(*)
    real(c_double), intent(out) :: divf(*)
    call lsb_check(lsb_div(lsb_ctx(), uf, vf, wf, divf), 'div')
  end subroutine div

  ! derivatives.f90:348-392
  subroutine divcheck(uf, vf, wf, diverr)
    real(c_double), intent(in) :: uf(*), vf(*), wf(*)
    real(c_double), intent(out) :: diverr
    call lsb_check(lsb_divcheck(lsb_ctx(), uf, vf, wf, diverr), 'divcheck')
  end subroutine divcheck

  ! derivatives.f90:394-409
  subroutine divuv(uf, vf, divf)
    real(c_double), intent(in) :: uf(*), vf(*)
    real(c_double), intent(out) :: divf(*)
    call lsb_check(lsb_divuv(lsb_ctx(), uf, vf, divf), 'divuv')
  end subroutine divuv
end module derivatives

! helmholtz.f90: eigen-decomposition and the Helmholtz / Poisson solves
module helmholtz
  use iso_c_binding
  use lsb_capi
  use lsb_context
  implicit none
  private
  public :: diagonalize_matrix, solve_helmholtz, solve_nhom_helmholtz
contains
  ! helmholtz.f90:13-126 (host)
  subroutine diagonalize_matrix(d2, n, lamda, e, ei)
    integer, intent(in) :: n
    real(c_double), intent(in) :: d2(*)
    real(c_double), intent(out) :: lamda(*), e(*), ei(*)
    call lsb_check(lsb_diagonalize_matrix(d2, int(n, c_int), lamda, e, ei), 'diagonalize_matrix')
  end subroutine diagonalize_matrix

  ! helmholtz.f90:130-208
  subroutine solve_helmholtz(rhs, dtnui, n, diag, smat, simat)
    integer, intent(in) :: n
    real(c_double), intent(inout) :: rhs(*)
    real(c_double), intent(in) :: dtnui
    real(c_double), intent(in) :: diag(*), smat(*), simat(*)
    call lsb_check(lsb_solve_helmholtz(lsb_ctx(), rhs, dtnui, int(n, c_int), diag, smat, simat), &
                   'solve_helmholtz')
  end subroutine solve_helmholtz

  ! helmholtz.f90:211-358
  subroutine solve_nhom_helmholtz(rhs, bc, dtnui, diag, smat, simat)
    real(c_double), intent(inout) :: rhs(*)
    real(c_double), intent(in) :: bc(*)
    real(c_double), intent(in) :: dtnui
    real(c_double), intent(in) :: diag(*), smat(*), simat(*)
    call lsb_check(lsb_solve_nhom_helmholtz(lsb_ctx(), rhs, bc, dtnui, diag, smat, simat), &
                   'solve_nhom_helmholtz')
  end subroutine solve_nhom_helmholtz
end module helmholtz

! sgsterm.f90: the subgrid-scale stress of the configured model
module sgs_models
  use iso_c_binding
  use lsb_capi
  use lsb_context
  implicit none
  private
  public :: tau_sgs, static_smagorinsky, dynamic_smagorinsky, sigma_model, interscale_model
contains
  ! sgsterm.f90:10-37
  subroutine tau_sgs(ui, uiuj, duidxj, tauij)
    real(c_double), intent(in) :: ui(*), uiuj(*), duidxj(*)
    real(c_double), intent(out) :: tauij(*)
    call lsb_check(lsb_tau_sgs(lsb_ctx(), ui, uiuj, duidxj, tauij), 'tau_sgs')
  end subroutine tau_sgs

  ! sgsterm.f90:39-120
  subroutine static_smagorinsky(cs, duidxj, tauij)
    real(c_double), intent(in) :: cs
    real(c_double), intent(in) :: duidxj(*)
    real(c_double), intent(out) :: tauij(*)
    call lsb_check(lsb_static_smagorinsky(lsb_ctx(), cs, duidxj, tauij), 'static_smagorinsky')
  end subroutine static_smagorinsky

  ! sgsterm.f90:123-280
  subroutine dynamic_smagorinsky(ui, uiuj, duidxj, tauij)
    real(c_double), intent(in) :: ui(*), uiuj(*), duidxj(*)
    real(c_double), intent(out) :: tauij(*)
    call lsb_check(lsb_dynamic_smagorinsky(lsb_ctx(), ui, uiuj, duidxj, tauij), 'dynamic_smagorinsky')
  end subroutine dynamic_smagorinsky

  ! sgsterm.f90:283-423
  subroutine sigma_model(duidxj, tauij)
    real(c_double), intent(in) :: duidxj(*)
    real(c_double), intent(out) :: tauij(*)
    call lsb_check(lsb_sigma_model(lsb_ctx(), duidxj, tauij), 'sigma_model')
  end subroutine sigma_model

  ! sgsterm.f90:426-467
  subroutine interscale_model(ui, tauij)
    real(c_double), intent(in) :: ui(*)
    real(c_double), intent(out) :: tauij(*)
    call lsb_check(lsb_interscale_model(lsb_ctx(), ui, tauij), 'interscale_model')
  end subroutine interscale_model
end module sgs_models

! filters.f90: the 3-point and spectral filters and the filtering of the state
module filters
  use iso_c_binding
  use lsb_capi
  use lsb_context
  implicit none
  private
  public :: filterall3pt, tns_automatic_filtering, filterqspec, specfilter_exp_xy, filter_q_3pt
contains
  ! filters.f90:9-49
  subroutine filterall3pt(uf, vf, wf, tempf, pf)
    real(c_double), intent(inout) :: uf(*), vf(*), wf(*), tempf(*), pf(*)
    call lsb_check(lsb_filterall3pt(lsb_ctx(), uf, vf, wf, tempf, pf), 'filterall3pt')
  end subroutine filterall3pt

  ! filters.f90:62-161
  subroutine tns_automatic_filtering(uf, vf, wf, tempf, pf)
    real(c_double), intent(inout) :: uf(*), vf(*), wf(*), tempf(*), pf(*)
    real(c_double) :: ih, idelta
    integer(c_int) :: applied
    call lsb_check(lsb_tns_automatic_filtering(lsb_ctx(), uf, vf, wf, tempf, pf, ih, idelta, applied), &
                   'tns_automatic_filtering')
  end subroutine tns_automatic_filtering

  ! filters.f90:164-192
  subroutine filterqspec(qf, nq)
    integer, intent(in) :: nq
    real(c_double), intent(inout) :: qf(*)
    call lsb_check(lsb_filterqspec(lsb_ctx(), qf, int(nq, c_int)), 'filterqspec')
  end subroutine filterqspec

  ! filters.f90:281-335
  subroutine specfilter_exp_xy(uf, uff, kc, p, amp, filtery)
    real(c_double), intent(in) :: uf(*)
    real(c_double), intent(out) :: uff(*)
    real(c_double), intent(in) :: kc, p, amp
    integer, intent(in) :: filtery
    call lsb_check(lsb_specfilter_exp_xy(lsb_ctx(), uf, uff, kc, p, amp, int(filtery, c_int)), 'specfilter_exp_xy')
  end subroutine specfilter_exp_xy

  ! filters.f90:363-458
  subroutine filter_q_3pt(q, q_filt, n4, n5, a, b, interpflag)
    integer, intent(in) :: n4, n5, interpflag
    real(c_double), intent(in) :: q(*)
    real(c_double), intent(out) :: q_filt(*)
    real(c_double), intent(in) :: a, b
    call lsb_check(lsb_filter_q_3pt(lsb_ctx(), q, q_filt, int(n4, c_int), int(n5, c_int), a, b, &
                                    int(interpflag, c_int)), 'filter_q_3pt')
  end subroutine filter_q_3pt
end module filters

! dealiasing.f90: the 3/2-rule products
module dealias
  use iso_c_binding
  use lsb_capi
  use lsb_context
  implicit none
  private
  public :: setup_dealiasing, dealiasing_2d_fftw, dealiasf_2d_fftw
  integer, public, save :: nx_lg = 0, ny_lg = 0
contains
  ! dealiasing.f90:87-188 (the padded sizes into nx_lg, ny_lg; the library keeps its own
  ! transforms, so ioption has nothing to select)
  subroutine setup_dealiasing(ioption)
    integer, intent(in) :: ioption
    integer(c_int) :: nxl, nyl
    call lsb_check(lsb_setup_dealiasing(lsb_ctx(), nxl, nyl), 'setup_dealiasing')
    nx_lg = nxl
    ny_lg = nyl
  end subroutine setup_dealiasing

  ! dealiasing.f90:262-447
  subroutine dealiasing_2d_fftw(a, b, ab)
    real(c_double), intent(in) :: a(*), b(*)
    real(c_double), intent(out) :: ab(*)
    call lsb_check(lsb_dealiasing_2d_fftw(lsb_ctx(), a, b, ab), 'dealiasing_2d_fftw')
  end subroutine dealiasing_2d_fftw

  ! dealiasing.f90:449-618
  subroutine dealiasf_2d_fftw(af, bf, abf)
    real(c_double), intent(in) :: af(*), bf(*)
    real(c_double), intent(out) :: abf(*)
    call lsb_check(lsb_dealiasf_2d_fftw(lsb_ctx(), af, bf, abf), 'dealiasf_2d_fftw')
  end subroutine dealiasf_2d_fftw
end module dealias
