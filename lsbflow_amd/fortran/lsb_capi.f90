! Fortran 2003 (iso_c_binding) interfaces to the C-ABI of liblsbflow_amd.so.
!
! One interface per entry point of include/lsbflow.h.  A driver that keeps the
! reference's module names (modhorfft, derivatives, helmholtz, sgs_models,
! filters, dealias, chebyshev) implements each module procedure as a thin call
! through these interfaces; arrays are passed by reference in the reference's
! own column-major layout, real(8) = real(c_double), integer = integer(c_int).
!
! Every function returns 0 (LSB_OK) or a negative status; lsb_error_message()
! turns the library's last message into a Fortran string for `error stop`.
module lsb_capi
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: LSB_OK = 0
  integer(c_int), parameter :: LSB_E_ARG = -1, LSB_E_HIP = -2, LSB_E_NOMEM = -3
  integer(c_int), parameter :: LSB_E_STATE = -4, LSB_E_UNSUPPORTED = -5, LSB_E_NUMERIC = -6
  integer(c_int), parameter :: LSB_FP64 = 0, LSB_MIXED = 1

  ! lsb_config: the module globals the hot path reads (common.f90:10-197)
  type, bind(C) :: lsb_config
    integer(c_int) :: nx, ny, nz
    integer(c_int) :: mapping, bctype, forcing
    integer(c_int) :: sgsflag, sgsmodel
    integer(c_int) :: dealiasflag, dealias_ref_compat
    integer(c_int) :: dct_exact, top_bc, dyn_avg, nut_ceil0
    integer(c_int) :: ifrz, nstart, precision, device
    real(c_double) :: xlen, ylen, zlen, x0, u0, xnu, blfac, prtl
    real(c_double) :: a, b, cs, kc, filt_order, filt_amp
    real(c_double) :: samp, istartbound
    real(c_double) :: dt
  end type lsb_config

  interface
    ! ------------------------------------------------------------ context
    integer(c_int) function lsb_create(cfg, ctx) bind(C, name="lsb_create")
      import :: c_int, c_ptr, lsb_config
      type(lsb_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function lsb_destroy(ctx) bind(C, name="lsb_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function lsb_last_error() bind(C, name="lsb_last_error")
      import :: c_ptr
    end function
    integer(c_int) function lsb_set_grid(ctx, wavx, wavy, zpts, gf, xpts, dx, dy, xl) &
        bind(C, name="lsb_set_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: wavx(*), wavy(*), zpts(*), gf(*), xpts(*)
      real(c_double), value :: dx, dy, xl
    end function
    integer(c_int) function lsb_set_mats(ctx, d, d2, diag, smat, simat) bind(C, name="lsb_set_mats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: d(*), d2(*), diag(*), smat(*), simat(*)
    end function
    integer(c_int) function lsb_set_paral(ctx, ubar, ubarp, ubarpp, wbar, wbarpp) &
        bind(C, name="lsb_set_paral")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ubar(*), ubarp(*), ubarpp(*), wbar(*), wbarpp(*)
    end function
    integer(c_int) function lsb_get_grid(ctx, wavx, wavy, zpts, gf, xpts, dxdyxl) bind(C, name="lsb_get_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: wavx(*), wavy(*), zpts(*), gf(*), xpts(*), dxdyxl(3)
    end function
    integer(c_int) function lsb_get_mats(ctx, dor, d, d2, mdz, md2, diag, smat, simat) &
        bind(C, name="lsb_get_mats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dor(*), d(*), d2(*), mdz(*), md2(*), diag(*), smat(*), simat(*)
    end function

    ! ------------------------------------------------- host-side setup
    integer(c_int) function lsb_build_cheb_zder_matrix(nzp, xl, gf, dor, d, d2) &
        bind(C, name="lsb_build_cheb_zder_matrix")
      import :: c_int, c_double
      integer(c_int), value :: nzp
      real(c_double), value :: xl
      real(c_double), intent(in) :: gf(*)
      real(c_double), intent(out) :: dor(*), d(*), d2(*)
    end function
    integer(c_int) function lsb_dct_diff(n1, n2, n, xl, gf, y, dydx, p) bind(C, name="lsb_dct_diff")
      import :: c_int, c_double
      integer(c_int), value :: n1, n2, n
      real(c_double), value :: xl
      real(c_double), intent(in) :: gf(*), y(*)
      real(c_double), intent(out) :: dydx(*)
      integer(c_int), value :: p
    end function
    integer(c_int) function lsb_diagonalize_matrix(d2, n, lamda, e, ei) bind(C, name="lsb_diagonalize_matrix")
      import :: c_int, c_double
      real(c_double), intent(in) :: d2(*)
      integer(c_int), value :: n
      real(c_double), intent(out) :: lamda(*), e(*), ei(*)
    end function

    ! ---------------------------------------------------------- modhorfft
    integer(c_int) function lsb_horfft(ctx, ain, ainf, is, nfields) bind(C, name="lsb_horfft")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: ain(*), ainf(*)
      integer(c_int), value :: is, nfields
    end function
    integer(c_int) function lsb_norm(ctx, uf, vf, wf, tempf) bind(C, name="lsb_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: uf(*), vf(*), wf(*), tempf(*)
    end function
    integer(c_int) function lsb_dataflip_xtoy(ctx, x, xf) bind(C, name="lsb_dataflip_xtoy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: xf(*)
    end function
    integer(c_int) function lsb_dataflip_ytox(ctx, x, xf) bind(C, name="lsb_dataflip_ytox")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: x(*)
      real(c_double), intent(in) :: xf(*)
    end function

    ! -------------------------------------------------------- derivatives
    integer(c_int) function lsb_dfdx(ctx, uf, dudxf) bind(C, name="lsb_dfdx")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: dudxf(*)
    end function
    integer(c_int) function lsb_dfdy(ctx, uf, dudyf) bind(C, name="lsb_dfdy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: dudyf(*)
    end function
    integer(c_int) function lsb_dfdz(ctx, uf, dudzf) bind(C, name="lsb_dfdz")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: dudzf(*)
    end function
    integer(c_int) function lsb_d2fdz2(ctx, uf, d2f) bind(C, name="lsb_d2fdz2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: d2f(*)
    end function
    integer(c_int) function lsb_dpdz(ctx, n1, n2, u, dudz, p) bind(C, name="lsb_dpdz")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1, n2
      real(c_double), intent(in) :: u(*)
      real(c_double), intent(out) :: dudz(*)
      integer(c_int), value :: p
    end function
    integer(c_int) function lsb_divuv(ctx, uf, vf, divf) bind(C, name="lsb_divuv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*)
      real(c_double), intent(out) :: divf(*)
    end function
    integer(c_int) function lsb_div(ctx, uf, vf, wf, divf) bind(C, name="lsb_div")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*)
      real(c_double), intent(out) :: divf(*)
    end function
    integer(c_int) function lsb_lapl(ctx, uf, lapluf) bind(C, name="lsb_lapl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: lapluf(*)
    end function
    integer(c_int) function lsb_divcheck(ctx, uf, vf, wf, diverr) bind(C, name="lsb_divcheck")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*)
      real(c_double), intent(out) :: diverr
    end function

    ! ---------------------------------------------------------- helmholtz
    integer(c_int) function lsb_solve_helmholtz(ctx, rhs, dtnui, n, diag, smat, simat) &
        bind(C, name="lsb_solve_helmholtz")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rhs(*)
      real(c_double), value :: dtnui
      integer(c_int), value :: n
      real(c_double), intent(in) :: diag(*), smat(*), simat(*)
    end function
    integer(c_int) function lsb_solve_nhom_helmholtz(ctx, rhs, bc, dtnui, diag, smat, simat) &
        bind(C, name="lsb_solve_nhom_helmholtz")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rhs(*)
      real(c_double), intent(in) :: bc(*)
      real(c_double), value :: dtnui
      real(c_double), intent(in) :: diag(*), smat(*), simat(*)
    end function

    ! --------------------------------------------------------- sgs_models
    integer(c_int) function lsb_tau_sgs(ctx, ui, uiuj, duidxj, tauij) bind(C, name="lsb_tau_sgs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ui(*), uiuj(*), duidxj(*)
      real(c_double), intent(out) :: tauij(*)
    end function
    integer(c_int) function lsb_static_smagorinsky(ctx, cs, duidxj, tauij) bind(C, name="lsb_static_smagorinsky")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cs
      real(c_double), intent(in) :: duidxj(*)
      real(c_double), intent(out) :: tauij(*)
    end function
    integer(c_int) function lsb_dynamic_smagorinsky(ctx, ui, uiuj, duidxj, tauij) &
        bind(C, name="lsb_dynamic_smagorinsky")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ui(*), uiuj(*), duidxj(*)
      real(c_double), intent(out) :: tauij(*)
    end function
    integer(c_int) function lsb_sigma_model(ctx, duidxj, tauij) bind(C, name="lsb_sigma_model")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: duidxj(*)
      real(c_double), intent(out) :: tauij(*)
    end function
    integer(c_int) function lsb_interscale_model(ctx, ui, tauij) bind(C, name="lsb_interscale_model")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ui(*)
      real(c_double), intent(out) :: tauij(*)
    end function

    ! ------------------------------------------------------------ filters
    integer(c_int) function lsb_filter_q_3pt(ctx, q, q_filt, n4, n5, a, b, interpflag) &
        bind(C, name="lsb_filter_q_3pt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*)
      real(c_double), intent(out) :: q_filt(*)
      integer(c_int), value :: n4, n5
      real(c_double), value :: a, b
      integer(c_int), value :: interpflag
    end function
    integer(c_int) function lsb_filterqspec(ctx, qf, nq) bind(C, name="lsb_filterqspec")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: qf(*)
      integer(c_int), value :: nq
    end function
    integer(c_int) function lsb_specfilter_exp_xy(ctx, uf, uff, kc, p, amp, filtery) &
        bind(C, name="lsb_specfilter_exp_xy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: uff(*)
      real(c_double), value :: kc, p, amp
      integer(c_int), value :: filtery
    end function

    ! ------------------------------------------------------------- dealias
    integer(c_int) function lsb_setup_dealiasing(ctx, nx_lg, ny_lg) bind(C, name="lsb_setup_dealiasing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nx_lg, ny_lg
    end function
    integer(c_int) function lsb_dealiasing_2d_fftw(ctx, a, b, ab) bind(C, name="lsb_dealiasing_2d_fftw")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: a(*), b(*)
      real(c_double), intent(out) :: ab(*)
    end function

    ! -------------------------------------------------------- step routines
    integer(c_int) function lsb_nlterm_skewsym(ctx, uif, nltermif) bind(C, name="lsb_nlterm_skewsym")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uif(*)
      real(c_double), intent(out) :: nltermif(*)
    end function
    integer(c_int) function lsb_state_set(ctx, uf, vf, wf, pf, pnf, unif, unmif, nlnif, nlnmif, istep, t) &
        bind(C, name="lsb_state_set")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*), pf(*), pnf(*)
      real(c_double), intent(in) :: unif(*), unmif(*), nlnif(*), nlnmif(*)
      integer(c_int), value :: istep
      real(c_double), value :: t
    end function
    integer(c_int) function lsb_state_get(ctx, uf, vf, wf, pf, pnf, unif, unmif, nlnif, nlnmif, istep, t) &
        bind(C, name="lsb_state_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: uf(*), vf(*), wf(*), pf(*), pnf(*)
      real(c_double), intent(out) :: unif(*), unmif(*), nlnif(*), nlnmif(*)
      integer(c_int), intent(out) :: istep
      real(c_double), intent(out) :: t
    end function
    integer(c_int) function lsb_timestep(ctx, nsteps) bind(C, name="lsb_timestep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function lsb_navierstokes_update(ctx) bind(C, name="lsb_navierstokes_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lsb_state_device_ptrs(ctx, uf, vf, wf, pf) bind(C, name="lsb_state_device_ptrs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: uf, vf, wf, pf
    end function
    type(c_ptr) function lsb_stream(ctx) bind(C, name="lsb_stream")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lsb_synchronize(ctx) bind(C, name="lsb_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! ------------------------------------------- device-pointer entry points
    integer(c_int) function lsb_horfft_dev(ctx, ain, ainf, is, nfields, stream) bind(C, name="lsb_horfft_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ain, ainf
      integer(c_int), value :: is, nfields
      type(c_ptr), value :: stream
    end function
    ! batched Helmholtz solves in FP64, FP32 or FP32 + one FP64 refinement (mode 0, 1, 2)
    integer(c_int) function lsb_solve_helmholtz_mixed_dev(ctx, rhs, dtnui, nfields, mode, stream) &
        bind(C, name="lsb_solve_helmholtz_mixed_dev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, rhs
      real(c_double), value :: dtnui
      integer(c_int), value :: nfields, mode
      type(c_ptr), value :: stream
    end function
    ! ensembles of contexts stepped from captured graphs
    integer(c_int) function lsb_ensemble_create(members, n, ens) bind(C, name="lsb_ensemble_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: members(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: ens
    end function
    integer(c_int) function lsb_ensemble_timestep(ens, nsteps) bind(C, name="lsb_ensemble_timestep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function lsb_ensemble_synchronize(ens) bind(C, name="lsb_ensemble_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function
    integer(c_int) function lsb_ensemble_info(ens, nmembers, ngraphs, phase) bind(C, name="lsb_ensemble_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), intent(out) :: nmembers, ngraphs, phase
    end function
    integer(c_int) function lsb_ensemble_destroy(ens) bind(C, name="lsb_ensemble_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function
    ! launch-shape autotuner
    integer(c_int) function lsb_autotune(ctx) bind(C, name="lsb_autotune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lsb_autotune_get(ctx, variant, ms, n) bind(C, name="lsb_autotune_get")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: variant(*)
      real(c_float), intent(out) :: ms(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function lsb_autotune_set(ctx, variant, n) bind(C, name="lsb_autotune_set")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: variant(*)
      integer(c_int), value :: n
    end function
    ! lsb_horfft_dev without device allocation (callers that cannot allocate)
    integer(c_int) function lsb_horfft_dev_noalloc(ctx, ain, ainf, is, nfields, stream) &
        bind(C, name="lsb_horfft_dev_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ain, ainf
      integer(c_int), value :: is, nfields
      type(c_ptr), value :: stream
    end function
    integer(c_int) function lsb_dfdz_dev(ctx, uf, dudzf, nfields, stream) bind(C, name="lsb_dfdz_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, uf, dudzf
      integer(c_int), value :: nfields
      type(c_ptr), value :: stream
    end function
    integer(c_int) function lsb_solve_helmholtz_dev(ctx, rhs, dtnui, nfields, stream) &
        bind(C, name="lsb_solve_helmholtz_dev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, rhs
      real(c_double), value :: dtnui
      integer(c_int), value :: nfields
      type(c_ptr), value :: stream
    end function

    ! ------------------------------------------------------------ profiling
    integer(c_int) function lsb_profile_begin(ctx) bind(C, name="lsb_profile_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lsb_profile_begin_classes(ctx, class_mask) bind(C, name="lsb_profile_begin_classes")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: class_mask
    end function
    integer(c_int) function lsb_profile_end(ctx, ms, count, flops, bytes) bind(C, name="lsb_profile_end")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(6)
      integer(c_int), intent(out) :: count(6)
      real(c_double), intent(out) :: flops(6), bytes(6)
    end function

    ! ------------------------------ snapshots, restart files, diagnostics
    ! path / rec arguments are type(c_ptr): c_loc of a NUL-terminated character
    ! array / real(c_float) buffer, or c_null_ptr
    integer(c_int) function lsb_snapshot_record(ctx, uf, vf, wf, tempf, pf, tbar, name, isum, t, dt, rec) &
        bind(C, name="lsb_snapshot_record")
      import :: c_int, c_ptr, c_double, c_float
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*), tempf(*), pf(*), tbar(*)
      integer(c_int), value :: name, isum
      real(c_double), value :: t, dt
      real(c_float), intent(out) :: rec(*)
    end function
    integer(c_int) function lsb_state_snapshot(ctx, name, path, rec) bind(C, name="lsb_state_snapshot")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: name
      type(c_ptr), value :: path, rec
    end function
    integer(c_int) function lsb_dealiasf_2d_fftw(ctx, af, bf, abf) bind(C, name="lsb_dealiasf_2d_fftw")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: af(*), bf(*)
      real(c_double), intent(out) :: abf(*)
    end function
    integer(c_int) function lsb_test_fft(ctx, u, avgerror) bind(C, name="lsb_test_fft")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
      real(c_double), intent(out) :: avgerror
    end function
    integer(c_int) function lsb_zinteg(ctx, tmp, tint, n1, n2, mapping, per, rev) bind(C, name="lsb_zinteg")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tmp(*)
      real(c_double), intent(out) :: tint(*)
      integer(c_int), value :: n1, n2, mapping, per, rev
    end function
    integer(c_int) function lsb_set_time_average(ctx, otfavg, istartavg, iomod) bind(C, name="lsb_set_time_average")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: otfavg, istartavg, iomod
    end function
    integer(c_int) function lsb_state_snapshot_average(ctx, path, rec) bind(C, name="lsb_state_snapshot_average")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: path, rec
    end function
    integer(c_int) function lsb_state_get_time_average(ctx, uf, vf, wf, tempf, pf) &
        bind(C, name="lsb_state_get_time_average")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: uf(*), vf(*), wf(*), tempf(*), pf(*)
    end function
    integer(c_int) function lsb_read_snapshot(ctx, path, q, isum, t, dt) bind(C, name="lsb_read_snapshot")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(out) :: q(*)
      integer(c_int), intent(out) :: isum
      real(c_double), intent(out) :: t, dt
    end function
    integer(c_int) function lsb_write_restart(ctx, path, rec) bind(C, name="lsb_write_restart")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, path, rec
    end function
    integer(c_int) function lsb_read_restart(ctx, path, rec) bind(C, name="lsb_read_restart")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, path, rec
    end function
    integer(c_int) function lsb_maxcalc(ctx, uf, vf, wf, tempf, tbar, dtmax, t, fsize, dt, dt2, dt3, stats, &
        change_dt) bind(C, name="lsb_maxcalc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*), tempf(*), tbar(*)
      real(c_double), value :: dtmax, t, fsize
      real(c_double), intent(inout) :: dt, dt2, dt3
      real(c_double), intent(out) :: stats(11)
      integer(c_int), intent(out) :: change_dt
    end function
    integer(c_int) function lsb_state_maxcalc(ctx, dtmax, fsize, stats, change_dt) bind(C, name="lsb_state_maxcalc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dtmax, fsize
      real(c_double), intent(out) :: stats(11)
      integer(c_int), intent(out) :: change_dt
    end function
    integer(c_int) function lsb_nancheck(ctx, uf, vf, wf, tempf, has_nan) bind(C, name="lsb_nancheck")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*), tempf(*)
      integer(c_int), intent(out) :: has_nan
    end function
    integer(c_int) function lsb_state_nancheck(ctx, has_nan) bind(C, name="lsb_state_nancheck")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: has_nan
    end function
    integer(c_int) function lsb_wall_shear(ctx, uf, cf, x_sep, x_reatt) bind(C, name="lsb_wall_shear")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*)
      real(c_double), intent(out) :: cf(*), x_sep, x_reatt
    end function
    integer(c_int) function lsb_state_wall_shear(ctx, cf, x_sep, x_reatt) bind(C, name="lsb_state_wall_shear")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: cf(*), x_sep, x_reatt
    end function
    integer(c_int) function lsb_sponge_lambda(nx, xpts, xlen, u0, fsize, lfsize, sfsize, lam) &
        bind(C, name="lsb_sponge_lambda")
      import :: c_int, c_double
      integer(c_int), value :: nx
      real(c_double), intent(in) :: xpts(*)
      real(c_double), value :: xlen, u0, fsize, lfsize, sfsize
      real(c_double), intent(out) :: lam(*)
    end function
    integer(c_int) function lsb_filterall3pt(ctx, uf, vf, wf, tempf, pf) bind(C, name="lsb_filterall3pt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: uf(*), vf(*), wf(*), tempf(*), pf(*)
    end function
    integer(c_int) function lsb_tns_automatic_filtering(ctx, uf, vf, wf, tempf, pf, ih, idelta, applied) &
        bind(C, name="lsb_tns_automatic_filtering")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: uf(*), vf(*), wf(*), tempf(*), pf(*)
      real(c_double), intent(out) :: ih, idelta
      integer(c_int), intent(out) :: applied
    end function
    integer(c_int) function lsb_state_filterall3pt(ctx) bind(C, name="lsb_state_filterall3pt")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lsb_state_tns_filter(ctx, ih, idelta, applied) bind(C, name="lsb_state_tns_filter")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ih, idelta
      integer(c_int), intent(out) :: applied
    end function
    integer(c_int) function lsb_ener(ctx, uf, vf, wf, itot, en) bind(C, name="lsb_ener")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*)
      integer(c_int), value :: itot
      real(c_double), intent(out) :: en
    end function
    integer(c_int) function lsb_lintest(ctx, uf, vf, wf, istep, dt, enr1, enr2, grow1, grow2) &
        bind(C, name="lsb_lintest")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uf(*), vf(*), wf(*)
      integer(c_int), value :: istep
      real(c_double), value :: dt
      real(c_double), intent(inout) :: enr1(*), enr2(*), grow1(*), grow2(*)
    end function
    integer(c_int) function lsb_state_lintest(ctx, enr1, enr2, grow1, grow2) bind(C, name="lsb_state_lintest")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: enr1(*), enr2(*), grow1(*), grow2(*)
    end function
    integer(c_int) function lsb_solve_orrsomm(nzp, d, d2, ubar, ubarp, ubarpp, zpts, xnu, eps, alphax, betax, &
        omega, iglb, uvec, vvec, wvec) bind(C, name="lsb_solve_orrsomm")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nzp, iglb
      real(c_double), intent(in) :: d(*), d2(*), ubar(*), ubarp(*), ubarpp(*), zpts(*)
      real(c_double), value :: xnu, eps, alphax, betax
      complex(c_double_complex), intent(inout) :: omega            ! double[2] (Re, Im) on the C side
      complex(c_double_complex), intent(out) :: uvec(*), vvec(*), wvec(*)
    end function
    integer(c_int) function lsb_set_temperature(ctx, tbar, tbarp, tbarpp, coex, gx, gy, gz) &
        bind(C, name="lsb_set_temperature")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tbar(*), tbarp(*), tbarpp(*)
      real(c_double), value :: coex, gx, gy, gz
    end function
    integer(c_int) function lsb_state_set_temperature(ctx, tempf, tnf, tnmf, nltnf, nltnmf) &
        bind(C, name="lsb_state_set_temperature")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tempf(*), tnf(*), tnmf(*), nltnf(*), nltnmf(*)
    end function
    integer(c_int) function lsb_state_get_temperature(ctx, tempf, tnf, tnmf, nltnf, nltnmf) &
        bind(C, name="lsb_state_get_temperature")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tempf(*), tnf(*), tnmf(*), nltnf(*), nltnmf(*)
    end function
    integer(c_int) function lsb_set_filtering(ctx, ifilt, modfilt) bind(C, name="lsb_set_filtering")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ifilt, modfilt
    end function
    integer(c_int) function lsb_set_sponge(ctx, fsize, lfsize, sfsize) bind(C, name="lsb_set_sponge")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: fsize, lfsize, sfsize
    end function
    integer(c_int) function lsb_set_adaptive_dt(ctx, dtadapt, iomod, fsize) bind(C, name="lsb_set_adaptive_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: dtadapt, iomod
      real(c_double), value :: fsize
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! The library's last error message as a Fortran string.
  function lsb_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = lsb_last_error()
    n = int(c_strlen(p))
    allocate(character(len=n) :: msg)
    if (n == 0) return
    call c_f_pointer(p, chars, [n])
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function lsb_error_message

  ! Status check used by the shim procedures: stop with the library's message.
  subroutine lsb_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= LSB_OK) then
      write(*, *) 'lsbflow_amd: ', where, ' failed (', status, '): ', lsb_error_message()
      error stop 1
    end if
  end subroutine lsb_check

end module lsb_capi
