/*
 * lsbflow.h -- C-ABI of the MI355X (gfx950) implementation of lsbflow's
 * per-time-step hot path.
 *
 * Every entry point below replaces one Fortran module procedure of the
 * reference (cited as file:line).  Arrays use the reference's own memory
 * layout (Fortran column-major, real(8)):
 *
 *   spectral field  qf(nxpl, ny, nzp)   nxpl = nx reals = nx/2 complex (Re,Im)
 *   physical field  q (nxpp, nypl, nzp) nxpp = nx + 2 (2 pad reals per x-row)
 *   vector          (..., 3)            three fields back to back
 *   tensor          (..., 3, 3)         element (l,m) is field l + 3*m
 *
 * with one rank (P = 1): nxpl = nx, nypl = ny, nzpl = nzp.
 *
 * Two tiers share one set of kernels:
 *   tier A  lsb_<name>(ctx, host pointers ...)   synchronous drop-in: copies
 *           in, runs the HIP kernels, copies out (correctness/parity path);
 *   tier B  lsb_state_* / lsb_timestep           the whole step device
 *           resident (throughput path), plus lsb_<name>_dev(ctx, device
 *           pointers ..., stream) kernel-level entry points.
 *
 * Errors: every function returns 0 on success or a negative LSB_E* code;
 * lsb_last_error() returns a message.  The reference has no return codes (it
 * prints and continues or calls MPI_FINALIZE, horfft.f90:327-330); a Fortran
 * shim turns a nonzero status into `error stop`.
 *
 * Pads of physical outputs are written as zero (the reference leaves FFTW
 * junk there, horfft.f90:144); pads of physical inputs are ignored, as in the
 * reference's r2c transform.
 */
#ifndef LSBFLOW_H
#define LSBFLOW_H

#ifdef __cplusplus
extern "C" {
#endif

#define LSB_OK 0
#define LSB_E_ARG (-1)
#define LSB_E_HIP (-2)
#define LSB_E_NOMEM (-3)
#define LSB_E_STATE (-4)
#define LSB_E_UNSUPPORTED (-5)
#define LSB_E_NUMERIC (-6)

#define LSB_FP64 0
#define LSB_MIXED 1

typedef struct lsb_ctx lsb_ctx;

/* Snapshot of the module globals the hot path reads through `use`
 * (common.f90:10-197: dim, runparam, parameters, sgsinput, time). */
typedef struct lsb_config {
    int nx, ny, nz;                 /* &dimen (common.f90:20)                  */
    int mapping;                    /* &run  mapping 0..3 (init.f90:188-254)    */
    int bctype;                     /* 0 homogeneous, 1 Spalart, 3 ceiling V(x) */
    int forcing;                    /* 0 periodic, 1 sponge layer (lsb_set_sponge) */
    int sgsflag, sgsmodel;          /* sgsterm.f90:21-35 (4: TNS, no SGS stress) */
    int dealiasflag;                /* 0 none, 1 3/2-rule, 2 spectral filter     */
    int dealias_ref_compat;         /* 1: reference ky placement (dealiasing.f90:329,435) */
    int dct_exact;                  /* 0: dct_diff as chebyshev.f90:242 (alpha(n) not
                                       scaled); 1: alpha(n)/xl, the exact Gauss-Lobatto
                                       derivative (identical when xl = 1)           */
    int top_bc;                     /* 0: u, v at the ceiling from D^2(u,v) -> 0
                                       (navierstokes.f90:1335-1347); 1: the wall's
                                       Zang-Hussaini values mirrored at the ceiling;
                                       2 (bctype 3): (1 - m(x)) free + m(x) Dirichlet,
                                       m = the profile's blowing Gaussian G(0.4 Lx) */
    int dyn_avg;                    /* dynamic model <M:L>/<M:M>: 0 the local 3x3x3 filter
                                       (sgsterm.f90:243-244), 1 the x-y plane means      */
    int nut_ceil0;                  /* 1: zero filter width (so zero eddy viscosity) on the
                                       ceiling row for the static (sgsterm.f90:95-103) and
                                       sigma (:368-376) models, as the wall row has (Q8);
                                       0: the reference's widths                         */
    int ifrz;                      /* navierstokes.f90:53-59                    */
    int nstart;                     /* 0: Euler/BDF2 start (navierstokes.f90:1507) */
    int precision;                  /* LSB_FP64 or LSB_MIXED                     */
    int device;                     /* HIP device ordinal                        */
    double xlen, ylen, zlen, x0, u0, xnu, blfac, prtl;  /* &param               */
    double a, b, cs, kc, filt_order, filt_amp;          /* &sgsin               */
    double samp, istartbound;       /* ceiling amplitude, BC start time          */
    double dt;                      /* constant time step (dtadapt = 0)          */
} lsb_config;

/* ---------------------------------------------------------------- context */

/* Builds dims (common.f90:23-54), grid (init.f90:130-273), Chebyshev operators
 * (chebyshev.f90:9-48, 204-270 as a matrix) and the FP64 eigen-setup
 * (helmholtz.f90:13-126), and uploads them.  Equivalent of the driver's
 * initialize() setup_fftw/setup_grid/setup_zder sequence (init.f90:36-97). */
int lsb_create(const lsb_config* cfg, lsb_ctx** out);
int lsb_destroy(lsb_ctx* ctx);
const char* lsb_last_error(void);

/* Override the snapshot with the driver's own module arrays (e.g. after
 * read_grid/read_paral on a restart, io.f90:134-316).  Lengths: wavx nx,
 * wavy ny, zpts/gf nzp, xpts nx; d/d2 nzp*nzp; diag nzm, smat/simat nzm*nzm. */
int lsb_set_grid(lsb_ctx* ctx, const double* wavx, const double* wavy, const double* zpts,
                 const double* gf, const double* xpts, double dx, double dy, double xl);
int lsb_set_mats(lsb_ctx* ctx, const double* d, const double* d2, const double* diag,
                 const double* smat, const double* simat);
/* paral module (common.f90:151-159): ubar, ubarp, ubarpp, wbar, wbarpp, length nzp. */
int lsb_set_paral(lsb_ctx* ctx, const double* ubar, const double* ubarp, const double* ubarpp,
                  const double* wbar, const double* wbarpp);
/* Host copies of the context's grid/operators (any pointer may be NULL). */
int lsb_get_grid(const lsb_ctx* ctx, double* wavx, double* wavy, double* zpts, double* gf,
                 double* xpts, double* dxdyxl /* [3] */);
int lsb_get_mats(const lsb_ctx* ctx, double* dor, double* d, double* d2, double* mdz,
                 double* md2, double* diag, double* smat, double* simat);

/* --------------------------------------------- host-side setup (no GPU) */

/* chebyshev.f90:9-48 build_cheb_zder_matrix(dor, d, d2): CGL matrix scaled 2/xl,
 * d = diag(gf) dor, d2 = d d.  Arrays nzp*nzp, column-major. */
int lsb_build_cheb_zder_matrix(int nzp, double xl, const double* gf, double* dor, double* d,
                               double* d2);
/* chebyshev.f90:204-270 dct_diff(n1, n2, n, y, dydx, p): DCT-I recurrence
 * derivative along the last (z) dimension, applied p times.  Host FP64. */
int lsb_dct_diff(int n1, int n2, int n, double xl, const double* gf, const double* y,
                 double* dydx, int p);
/* helmholtz.f90:13-126 diagonalize_matrix(D2, n, lamda, e, ei): D2 = E diag(lamda) E^-1
 * (own Hessenberg + Francis QR + inverse iteration in extended precision;
 * replaces LAPACK dggev/dgetrf/dgetri).  Column-major n*n. */
int lsb_diagonalize_matrix(const double* D2, int n, double* lamda, double* e, double* ei);

/* ------------------------------------------------------ tier A: modhorfft */

/* horfft.f90:61-150 horfft(ain, ainf, is): is = -1 physical -> spectral (scaled
 * 1/(nx*ny), kx-Nyquist dropped); is = +1 spectral -> physical.  nfields
 * consecutive fields (the reference calls it per field). */
int lsb_horfft(lsb_ctx* ctx, double* ain, double* ainf, int is, int nfields);
/* horfft.f90:153-224 norm(uf, vf, wf, tempf). */
int lsb_norm(lsb_ctx* ctx, double* uf, double* vf, double* wf, double* tempf);
/* horfft.f90:261-356 / 359-454 at P = 1: the layout reorder between the
 * x-row (nxpp) and spectral (nxpl) strides. */
int lsb_dataflip_xtoy(lsb_ctx* ctx, const double* x, double* xf);
int lsb_dataflip_ytox(lsb_ctx* ctx, double* x, const double* xf);

/* --------------------------------------------------- tier A: derivatives */

int lsb_dfdx(lsb_ctx* ctx, const double* uf, double* dudxf);      /* derivatives.f90:261-280 */
int lsb_dfdy(lsb_ctx* ctx, const double* uf, double* dudyf);      /* derivatives.f90:283-303 */
int lsb_dfdz(lsb_ctx* ctx, const double* uf, double* dudzf);      /* derivatives.f90:112-132 */
int lsb_d2fdz2(lsb_ctx* ctx, const double* uf, double* d2f);      /* derivatives.f90:161-188 */
/* derivatives.f90:135-158 dpdz(n1, n2, n, u, dudz, p) with n = nzp (device). */
int lsb_dpdz(lsb_ctx* ctx, int n1, int n2, const double* u, double* dudz, int p);
int lsb_divuv(lsb_ctx* ctx, const double* uf, const double* vf, double* divf); /* :394-409 */
int lsb_div(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf,
            double* divf);                                          /* :329-346 */
int lsb_lapl(lsb_ctx* ctx, const double* uf, double* lapluf);     /* :306-326 */
/* derivatives.f90:348-392 divcheck: mean |div u| in physical space. */
int lsb_divcheck(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf,
                 double* diverr);

/* ----------------------------------------------------- tier A: helmholtz */

/* helmholtz.f90:130-208 solve_helmholtz(rhs, dtnui, n, diag, smat, simat), n = nzm. */
int lsb_solve_helmholtz(lsb_ctx* ctx, double* rhs, double dtnui, int n, const double* diag,
                        const double* smat, const double* simat);
/* helmholtz.f90:211-358 solve_nhom_helmholtz(rhs, bc, dtnui, diag, smat, simat). */
int lsb_solve_nhom_helmholtz(lsb_ctx* ctx, double* rhs, const double* bc, double dtnui,
                             const double* diag, const double* smat, const double* simat);

/* ---------------------------------------------------- tier A: sgs_models */

/* sgsterm.f90:10-37 tau_sgs(ui, uiuj, duidxj, tauij), physical (nxpp,ny,nzp,3[,3]). */
int lsb_tau_sgs(lsb_ctx* ctx, const double* ui, const double* uiuj, const double* duidxj,
                double* tauij);
int lsb_static_smagorinsky(lsb_ctx* ctx, double cs, const double* duidxj, double* tauij); /* :39-120 */
int lsb_dynamic_smagorinsky(lsb_ctx* ctx, const double* ui, const double* uiuj,
                            const double* duidxj, double* tauij);                     /* :123-280 */
int lsb_sigma_model(lsb_ctx* ctx, const double* duidxj, double* tauij);               /* :283-423 */
int lsb_interscale_model(lsb_ctx* ctx, const double* ui, double* tauij);              /* :426-467 */

/* ------------------------------------------------------- tier A: filters */

/* filters.f90:363-458 filter_q_3pt(q, q_filt, n4, n5, a, b, interpflag). */
int lsb_filter_q_3pt(lsb_ctx* ctx, const double* q, double* q_filt, int n4, int n5, double a,
                     double b, int interpflag);
/* filters.f90:164-192 filterqspec(qf, nq) (in place). */
int lsb_filterqspec(lsb_ctx* ctx, double* qf, int nq);
/* filters.f90:281-335 specfilter_exp_xy(uf, uff, kc, p, amp, filtery). */
int lsb_specfilter_exp_xy(lsb_ctx* ctx, const double* uf, double* uff, double kc, double p,
                          double amp, int filtery);

/* ------------------------------------------------------- tier A: dealias */

/* dealiasing.f90:87-188 setup_dealiasing: returns nx_lg, ny_lg. */
int lsb_setup_dealiasing(lsb_ctx* ctx, int* nx_lg, int* ny_lg);
/* dealiasing.f90:262-447 dealiasing_2d_fftw(a, b, ab): padded product of two
 * physical fields (ref_compat placement per lsb_config.dealias_ref_compat). */
int lsb_dealiasing_2d_fftw(lsb_ctx* ctx, const double* a, const double* b, double* ab);
/* dealiasing.f90:449-618 dealiasf_2d_fftw(af, bf, abf): the same padded product of two
 * spectral fields (nxpl, ny, nzp), spectral out; the reference's ky placement as above. */
int lsb_dealiasf_2d_fftw(lsb_ctx* ctx, const double* af, const double* bf, double* abf);
/* horfft.f90:7-58 test_fft(u, avgerror): forward and inverse 2-D transform of the physical
 * field u (nxpp, ny, nzp) on the device; avgerror = mean |u - q| over the nx physical
 * points of every row. */
int lsb_test_fft(lsb_ctx* ctx, const double* u, double* avgerror);
/* derivatives.f90:9-70 zinteg(tmp, tint, n1, n2, mapping, per, rev): cumulative
 * wall-normal integral of every pencil of tmp (n1, n2, nzp) on the device -- per = 0 the
 * trapezoid rule on zpts (rev = 0 from the top plane nzp down, rev = 1 from plane 1 up),
 * per = 1 the interior quadrature weights with the change of variable 0.5 xl / gf;
 * mapping = 1 copies plane 2 to plane 1 (rev = 0 / per = 1). */
int lsb_zinteg(lsb_ctx* ctx, const double* tmp, double* tint, int n1, int n2, int mapping, int per, int rev);

/* ------------------------------------------------- tier A: step routines */

/* navierstokes.f90:658-855 nlterm_skewsym(uif, tempf, nltermif, nltempf, sgsflag)
 * for tempf == 0 (the temperature path is not restated; see lsb_state_set). */
int lsb_nlterm_skewsym(lsb_ctx* ctx, const double* uif, double* nltermif);

/* ----------------------------------------------------- tier B: the step */

/* Core state (common.f90:162-175: uf, vf, wf, pf, pnf, unif, unmif, nlnif,
 * nlnmif) and time module values; any history pointer may be NULL (zero).
 * istep sets both step counters: the run's own (start-up coefficients, adaptive-dt
 * schedule) and the total (isum of snapshot/restart headers); state_get returns the
 * total.  The step coefficients are recomputed on the first step after a set.
 * The temperature state is set apart (lsb_state_set_temperature). */
int lsb_state_set(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf,
                  const double* pf, const double* pnf, const double* unif, const double* unmif,
                  const double* nlnif, const double* nlnmif, int istep, double t);
int lsb_state_get(lsb_ctx* ctx, double* uf, double* vf, double* wf, double* pf, double* pnf,
                  double* unif, double* unmif, double* nlnif, double* nlnmif, int* istep,
                  double* t);
/* navierstokes.f90:10-90 timestep() x nsteps (ifilt = 0; dtadapt per lsb_set_adaptive_dt):
 * navierstokes_update + norm + ifrz + time update, all on the device. */
int lsb_timestep(lsb_ctx* ctx, int nsteps);
/* navierstokes.f90:1157-1496 navierstokes_update (one step, no norm). */
int lsb_navierstokes_update(lsb_ctx* ctx);
/* Temperature / buoyancy path (FP64 form; navierstokes.f90:800-845, 1251-1303): the paral
 * profiles tbar, tbarp, tbarpp (nzp each, NULL = 0) and the buoyancy coex, (gx, gy, gz).
 * Switches the path on. */
int lsb_set_temperature(lsb_ctx* ctx, const double* tbar, const double* tbarp, const double* tbarpp,
                        double coex, double gx, double gy, double gz);
/* core temperature state tempf, tnf, tnmf, nltnf, nltnmf (spectral, NULL = 0); switches the
 * path on.  The get returns zeros while the path is off. */
int lsb_state_set_temperature(lsb_ctx* ctx, const double* tempf, const double* tnf, const double* tnmf,
                              const double* nltnf, const double* nltnmf);
int lsb_state_get_temperature(lsb_ctx* ctx, double* tempf, double* tnf, double* tnmf, double* nltnf,
                              double* nltnmf);
/* Device pointers of the resident state (for in-HBM benchmarks): fields
 * uf, vf, wf, pf (each nx*ny*nzp doubles). */
int lsb_state_device_ptrs(lsb_ctx* ctx, void** uf, void** vf, void** wf, void** pf);
/* Stream the context launches on (hipStream_t). */
void* lsb_stream(lsb_ctx* ctx);
/* Wait for all work queued on the context stream (lsb_timestep only enqueues). */
int lsb_synchronize(lsb_ctx* ctx);

/* Launch-shape autotuner (no counterpart in the reference, whose FFTW plans are
 * chosen by FFTW_MEASURE, modfftw.f90:93-132): times every launch variant of each
 * tunable kernel family on this context's grid and keeps the fastest; scratch is
 * allocated for the duration of the call.  Families (slot: meaning): 0 the 3-point
 * filter on FP64 fields, 1 on FP32 fields (launch variant index); the kx columns per
 * tile (8, 16, or 32 with FP32 arithmetic) of 2 the forward y pass in FP64, 3 in FP32,
 * 4 the gradient transform's inverse y pass in FP64, 5 in FP32, 6 the mixed form's
 * forward y pass fused with the right-hand side; -1 in a slot means the built-in default.
 * lsb_autotune_get reads n slots (variant index, measured ms per launch),
 * lsb_autotune_set installs a saved table. */
int lsb_autotune(lsb_ctx* ctx);

/* Ensembles of independent contexts stepped from captured HIP graphs (no counterpart in
 * the reference, whose driver steps one domain, navierstokes.f90:10-90).  The members
 * (contexts on one device, each with its state set and past its three start-up steps,
 * constant dt, no adaptive dt / filtering / time averaging / profiling / temperature,
 * bctype 0 or 3) are captured together: one graph per phase of the history rotation
 * (60 steps), each graph one time step of every member on its own stream, the members
 * forked from and joined into the ensemble's stream.  lsb_ensemble_timestep launches
 * one graph per step and advances every member's step counter, time and history slots;
 * member state is read with lsb_state_get after lsb_ensemble_synchronize.  While in an
 * ensemble a member refuses lsb_timestep and may not grow its staging buffers. */
typedef struct lsb_ensemble lsb_ensemble;
int lsb_ensemble_create(lsb_ctx** members, int n, lsb_ensemble** out);
int lsb_ensemble_timestep(lsb_ensemble* ens, int nsteps);
int lsb_ensemble_synchronize(lsb_ensemble* ens);
int lsb_ensemble_info(const lsb_ensemble* ens, int* nmembers, int* ngraphs, int* phase);
int lsb_ensemble_destroy(lsb_ensemble* ens);
int lsb_autotune_get(const lsb_ctx* ctx, int* variant, float* ms, int n);
int lsb_autotune_set(lsb_ctx* ctx, const int* variant, int n);

/* ------------------------- snapshots, restart files, per-step diagnostics */

/* mpi_io.f90:339-497 mpi_snapshot(uf, vf, wf, tempf, pf, name) as the record it writes:
 * rec = 7 floats [isum, t, dt, nx, ny, nzp, nypl] then u, v, w, temp, p in physical
 * space, float32, x fastest then z then y (5*nx*ny*nzp floats).  u += ubar and
 * temp += tbar unless name == 2.  tempf and tbar may be NULL (zero).  With one rank
 * nypl = ny; the body is the same for any rank count. */
int lsb_snapshot_record(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf,
                        const double* tempf, const double* pf, const double* tbar, int name,
                        int isum, double t, double dt, float* rec);
/* The same record from the resident state (temp from the temperature path, else 0; tbar
 * of lsb_set_temperature; isum, t, dt of the context).  path != NULL writes the file (an
 * existing one is renamed prev_<name>, mpi_io.f90:400-404; the caller names it q%05d.dat /
 * qavg%05d.dat / q_ref.dat as :396-402); rec may be NULL. */
int lsb_state_snapshot(lsb_ctx* ctx, int name, const char* path, float* rec);
/* main.f90:69-76, 98-106 on-the-fly time averaging (runparam otfavg, istartavg, iomod):
 * with otfavg = 1, after every step of lsb_timestep whose total step count exceeds
 * istartavg, q_tavg = q_tavg + q/iomod for uf, vf, wf, tempf, pf, on the device.  The
 * averages start at zero here, on lsb_state_set and on lsb_read_restart (init.f90:112-119);
 * otfavg = 0 releases them. */
int lsb_set_time_average(lsb_ctx* ctx, int otfavg, int istartavg, int iomod);
/* mpi_snapshot(uf_tavg, vf_tavg, wf_tavg, tempf_tavg, pf_tavg, 1) then the averages back to
 * zero (main.f90:98-106): the qavg record (u + ubar, temp + tbar), path / rec as
 * lsb_state_snapshot.  Averaging off -> LSB_E_STATE. */
int lsb_state_snapshot_average(lsb_ctx* ctx, const char* path, float* rec);
/* The current averages, spectral (nxpl, ny, nzp) each; any pointer may be NULL. */
int lsb_state_get_time_average(lsb_ctx* ctx, double* uf, double* vf, double* wf, double* tempf, double* pf);
/* mpi_io.f90:500-650 mpi_read_snapshot: q(nx, ny, nzp, 5) physical (Fortran order) from a
 * snapshot file; dimension mismatch -> LSB_E_ARG.  Host-only. */
int lsb_read_snapshot(lsb_ctx* ctx, const char* path, double* q, int* isum, double* t, double* dt);
/* mpi_io.f90:9-160 mpi_write_restart_file: 9 floats [isum, t, dt, dt2, dt3, nx, ny, nzp,
 * nxpl] then 20 float32 spectral fields (uf, vf, wf, tempf, unif(3), unmif(3), nlnif(3),
 * nlnmif(3), tnf, tnmf, nltnf, nltnmf), z fastest then y then x.  Slots 14-16 hold
 * nlnmif (the reference writes nlnif twice there, quirk Q12).  Temperature slots hold
 * the temperature path's fields (0 while it is off).  From the resident state; path and/or
 * rec (9 + 20*nx*ny*nzp floats) may be NULL. */
int lsb_write_restart(lsb_ctx* ctx, const char* path, float* rec);
/* mpi_io.f90:163-336 mpi_read_restart_file into the resident state (rec, else the file):
 * the total step count = nint(isum), t, dt, dt2, dt3 from the header; the run's own step
 * counter starts again at 0 (main.f90:61), so the next step follows the nstart start-up
 * rule.  pf/pnf keep their values (zero on a fresh context).  Nonzero temperature
 * slots switch the temperature path on (FP64 form; in the mixed form -> LSB_E_UNSUPPORTED). */
int lsb_read_restart(lsb_ctx* ctx, const char* path, const float* rec);

/* navierstokes.f90:358-511 maxcalc(uf, vf, wf, tempf, dtmax, dt, dt2, dt3) with the hidden
 * inputs t (time), fsize (fringe size) and tbar (may be NULL).  stats[11] = umax, umin,
 * vmax, vmin, wmax, wmin, tmax, tmin (u + ubar, temp + tbar), x-, y-, z-CFL.  On a CFL
 * outside the limits dt2 = dt and dt = dt/1.25 (or dt*1.25); *change_dt = 1 then.
 * Extrema run over the nx physical points of each row (the reference also reads the
 * pad junk, quirk Q10). */
int lsb_maxcalc(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf,
                const double* tempf, const double* tbar, double dtmax, double t, double fsize,
                double* dt, double* dt2, double* dt3, double* stats, int* change_dt);
/* maxcalc on the resident state; updates the context's dt and dt2. */
int lsb_state_maxcalc(lsb_ctx* ctx, double dtmax, double fsize, double* stats, int* change_dt);
/* navierstokes.f90:77-81 / main.f90:79: any(isnan(.)) over the spectral fields (NULL
 * fields skipped). */
int lsb_nancheck(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf,
                 const double* tempf, int* has_nan);
int lsb_state_nancheck(lsb_ctx* ctx, int* has_nan);
/* Skin friction cf(nx) of run/compare_results.m:63,76-79 (spanwise mean of u + ubar,
 * xnu (u(nz) - u(nzp))/(zpts(nz) - zpts(nzp)) / (0.5 u0^2)) and the first zero crossings
 * of cf from >= 0 to < 0 (x_sep) and back (x_reatt), linear in xpts; NaN if none. */
int lsb_wall_shear(lsb_ctx* ctx, const double* uf, double* cf, double* x_sep, double* x_reatt);
int lsb_state_wall_shear(lsb_ctx* ctx, double* cf, double* x_sep, double* x_reatt);
/* navierstokes.f90:119-187 force(): the sponge coefficient lambda(x) (nx values) of the
 * fringe from xpts, xlen, u0, fsize, lfsize, sfsize.  Host only. */
int lsb_sponge_lambda(int nx, const double* xpts, double xlen, double u0, double fsize, double lfsize,
                      double sfsize, double* lam);
/* runparam fsize, lfsize, sfsize of the sponge layer (config forcing = 1; lsb_create uses
 * run/input.dat's 0.15, 0.25, 0.70 until this is called).  With forcing = 1 the step adds
 * lambda(x) (ubar - u, -v, -w) to the nonlinear term (navierstokes.f90:750-756) and the
 * velocity solves take their boundary values (navierstokes.f90:1389). */
int lsb_set_sponge(lsb_ctx* ctx, double fsize, double lfsize, double sfsize);
/* filters.f90:9-49 filterall3pt(uf, vf, wf, tempf, pf), in place: modecutFC (six 3-pt
 * filter passes with the SGS pair (a, b), coefficients 6, -15, 20, -15, 6, -1) of the
 * physical u, v, w, p, and temp when some |T| > 1e-10.  tempf may be NULL. */
int lsb_filterall3pt(lsb_ctx* ctx, double* uf, double* vf, double* wf, double* tempf, double* pf);
/* filters.f90:52-161 tns_automatic_filtering(uf, vf, wf, tempf, pf), in place: modecutFC with
 * (1, 6) and (1, 2); when the ratio of the kinetic energies they remove (trapezoidal z
 * integrals, I(h)/I(delta)) exceeds 0.007 the (1, 6) result replaces u, v, w, p.  ih,
 * idelta, applied (may be NULL) return the integrals and the decision.  tempf is not
 * touched (as the reference). */
int lsb_tns_automatic_filtering(lsb_ctx* ctx, double* uf, double* vf, double* wf, double* tempf,
                                double* pf, double* ih, double* idelta, int* applied);
/* The same two on the resident state (u, v, w, pf). */
int lsb_state_filterall3pt(lsb_ctx* ctx);
int lsb_state_tns_filter(lsb_ctx* ctx, double* ih, double* idelta, int* applied);
/* runparam ifilt, modfilt for lsb_timestep (navierstokes.f90:31-39): with ifilt >= 1, before
 * every step whose run counter is a multiple of modfilt, tns_automatic_filtering when
 * sgsflag = 1 and sgsmodel = 4, else filterall3pt. */
int lsb_set_filtering(lsb_ctx* ctx, int ifilt, int modfilt);
/* orrsomm.f90:625-666 ener(uf, vf, wf, itot, en): the kinetic energy 0.5 (u^2 + v^2 + w^2)
 * summed over each plane's physical points (u + ubar when itot = 1) and integrated in z by
 * the trapezoidal rule from row nz+1 up to row 2 (the top interval left out, k0 = 2). */
int lsb_ener(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf, int itot, double* en);
/* orrsomm.f90:566-622 lintest(uf, vf, wf, enr1, enr2, grow1, grow2) for step istep (1-based,
 * the main loop's counter) with the time module's dt: enr1 = ener(total), enr2 = ener with
 * the (0, 0) mode removed; grow = (E(istep) - E(istep-1))/(2 E(istep) dt), 0 at istep 1.
 * The arrays hold at least istep values; entry istep-1 is written. */
int lsb_lintest(lsb_ctx* ctx, const double* uf, const double* vf, const double* wf, int istep, double dt,
                double* enr1, double* enr2, double* grow1, double* grow2);
/* lintest on the resident state after lsb_timestep (the run's step counter, the context's dt). */
int lsb_state_lintest(lsb_ctx* ctx, double* enr1, double* enr2, double* grow1, double* grow2);
/* orrsomm.f90:110-562 solve_orrsomm(uvec, vvec, wvec, eps, alphax, betax, omega, iglb) from the
 * module arrays it reads: d, d2 (nzp x nzp column-major, mats), ubar, ubarp, ubarpp (paral),
 * zpts, xnu.  omega[2] (Re, Im) is the guess in and the eigenvalue out; uvec/vvec/wvec get
 * nzp complex values each (Re, Im interleaved), the mode scaled to energy eps.  Host only.
 * iglb = 0 (the iteration from the guess alpha*cguess*u0) only: iglb < 0 (a global zggevx
 * spectrum first; init.f90:606 passes -1) and iglb > 0 (no solve) -> LSB_E_UNSUPPORTED.  A failed iteration, which the reference answers
 * with a random restart, -> LSB_E_NUMERIC. */
int lsb_solve_orrsomm(int nzp, const double* d, const double* d2, const double* ubar, const double* ubarp,
                      const double* ubarpp, const double* zpts, double xnu, double eps, double alphax,
                      double betax, double* omega, int iglb, double* uvec, double* vvec, double* wvec);
/* runparam dtadapt, iomod, fsize for lsb_timestep: with dtadapt = 1, after every step
 * whose count is a multiple of 5 or of iomod, a NaN check (LSB_E_NUMERIC) and
 * maxcalc(dtmax = 1.0) on the state (navierstokes.f90:74-87). */
int lsb_set_adaptive_dt(lsb_ctx* ctx, int dtadapt, int iomod, double fsize);

/* ------------------------------------- device-pointer kernel entry points */
/* Same semantics as the tier-A routines, on device memory, asynchronous on
 * `stream` (NULL = the context stream). */
int lsb_horfft_dev(lsb_ctx* ctx, double* ain, double* ainf, int is, int nfields, void* stream);
int lsb_dfdz_dev(lsb_ctx* ctx, const double* uf, double* dudzf, int nfields, void* stream);
int lsb_solve_helmholtz_dev(lsb_ctx* ctx, double* rhs, double dtnui, int nfields, void* stream);

/* helmholtz.f90:130-208 (homogeneous Dirichlet rows) on nfields device spectral fields in
 * one of four precisions: LSB_HELM_FP64 the FP64 eigen-projections (as
 * lsb_solve_helmholtz_dev); LSB_HELM_FP32 both projections on the FP32 matrix cores;
 * LSB_HELM_REFINE the FP32 solve followed by one FP64 refinement (residual of
 * (D2 - k^2 - sigma) u = rhs in FP64 with D2 = d.d, FP32 solve of it, u += du);
 * LSB_HELM_SPLIT32 the mixed form's velocity solve: the FP64 leading term -rhs/(k^2 + sigma)
 * plus the FP32 eigen-projection correction, when sigma >= 16 max|lambda| (the FP64
 * projections otherwise). */
#define LSB_HELM_FP64 0
#define LSB_HELM_FP32 1
#define LSB_HELM_REFINE 2
#define LSB_HELM_SPLIT32 3
int lsb_solve_helmholtz_mixed_dev(lsb_ctx* ctx, double* rhs, double dtnui, int nfields, int mode, void* stream);

/* lsb_horfft_dev without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * lsb_horfft_dev; the inverse transform's staging buffer is a library-owned device
 * array of 256 MiB per GPU (2^25 doubles; resolved by lsb_create, held until the
 * process ends) instead of a buffer allocated on first use, and the call makes no
 * hipMalloc/hipFree.  LSB_E_ARG when one field's staging (nzp*ny*nx doubles) exceeds
 * the array; fields are transformed in batches that fit it.  All calls on one GPU
 * share the array: calls that overlap on a GPU (two streams, two threads) race. */
int lsb_horfft_dev_noalloc(lsb_ctx* ctx, double* ain, double* ainf, int is, int nfields, void* stream);

/* Kernel timing for roofline accounting.  Between lsb_profile_begin and
 * lsb_profile_end every launch group is bracketed by HIP events on the launch
 * stream and tagged with its algorithmic work.  Per class c (0 FFT, 1 Chebyshev
 * contraction, 2 Helmholtz, 3 other pointwise, 4 right-hand-side assembly,
 * 5 physical-space products and SGS model): total device ms, launch count,
 * algorithmic flops and algorithmic HBM bytes (arrays of LSB_PROF_CLASSES). */
#define LSB_PROF_CLASSES 6
int lsb_profile_begin(lsb_ctx* ctx);
/* The same, bracketing only the launch groups of the classes whose bit is set in
 * class_mask (bit c for class c); the others run without events.  No counterpart in the
 * reference: it keeps the timing events of a measured run to the class being measured. */
int lsb_profile_begin_classes(lsb_ctx* ctx, unsigned class_mask);
int lsb_profile_end(lsb_ctx* ctx, double* ms, int* count, double* flops, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* LSBFLOW_H */
