// Internal declarations of the lsbflow MI355X library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lsbflow.h"

namespace lsb {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);

#define LSB_HIP(call)                                                               \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::lsb::set_error(std::string(#call) + ": " + hipGetErrorString(e_));    \
            return LSB_E_HIP;                                                       \
        }                                                                           \
    } while (0)

#define LSB_TRY(call)                \
    do {                             \
        int rc_ = (call);            \
        if (rc_ != LSB_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------ complex math
template <class R>
struct cpx {
    R re, im;
};

template <class R>
__host__ __device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__host__ __device__ inline cpx<R> cadd(cpx<R> a, cpx<R> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class R>
__host__ __device__ inline cpx<R> csub(cpx<R> a, cpx<R> b) {
    return {a.re - b.re, a.im - b.im};
}
template <class R>
__host__ __device__ inline cpx<R> cconj(cpx<R> a) {
    return {a.re, -a.im};
}

// ------------------------------------------------------------- FFT plans
// A 1-D complex FFT of length L = prod(radix[0..nstage-1]) executed as a
// Stockham autosort sequence in LDS.  Twiddles are exp(-2 pi i m / L),
// m = 0..L-1, computed in long double on the host.
constexpr int kMaxStages = 16;
struct FftPlan1D {
    int L = 0;
    int nstage = 0;
    int radix[kMaxStages] = {0};
};

bool make_plan(int L, FftPlan1D* plan);

// Device-side description of one horizontal transform size.
struct HorPlan {
    int nx = 0, ny = 0;        // real length in x, complex length in y
    int lh = 0;                // nx/2 complex points per x-row
    FftPlan1D px, py;          // plans for lh and ny
    double* tw_x = nullptr;    // lh twiddles exp(-2 pi i m/lh) (cpx<double>)
    double* tw_y = nullptr;    // ny twiddles
    double* tw_r = nullptr;    // lh post-process twiddles exp(-2 pi i k/nx)
    float* tw_x32 = nullptr;
    float* tw_y32 = nullptr;
    float* tw_r32 = nullptr;
};

// A list of field base pointers passed by value to batched kernels.
// A z operator split by reflection parity for a symmetric grid (zops.hip):
// P and Q, me x me each, column-major with leading dimension ldm, Q at d_m + qoff.
struct ZSplitOp {
    bool ok = false;
    int n = 0, me = 0, ldm = 0;
    double osgn = 1.0;
    double* d_m = nullptr;
    float* d_m32 = nullptr;   // the same block in FP32 (FP32 split kernel)
    size_t qoff = 0;
};

constexpr int kMaxFields = 16;
// launch-shape tuning table of a context (lsb_autotune): the variant index per kernel family
// (the y-pass families hold the kx columns per tile: 8, 16 or, FP32 only, 32)
enum {
    TUNE_FILT3_F64 = 0,
    TUNE_FILT3_F32 = 1,
    TUNE_YFWD_F64 = 2,
    TUNE_YFWD_F32 = 3,
    TUNE_YINV3_F64 = 4,
    TUNE_YINV3_F32 = 5,
    TUNE_YRHS_F32 = 6,
    TUNE_SLOTS = 16
};
// leading dimension of the two-row boundary operator d_md2b (rows 0 and nzp-1 of d2fdz2)
constexpr int kLdmBnd = 64;
// doubles in the reduction / small-table scratch d_red
constexpr int kRedDoubles = 8192;
// doubles in the scratch of the diagnostics reductions (d_diag_red)
constexpr int kDiagDoubles = 16384;
template <class R>
struct FieldList {
    R* p[kMaxFields];
};

}  // namespace lsb

// --------------------------------------------------------------- context
struct lsb_ctx {
    lsb_config cfg{};
    int nx = 0, ny = 0, nz = 0, nzp = 0, nzm = 0, nxh = 0, nyh = 0, nxpp = 0;
    size_t nspec = 0;   // nx*ny*nzp   reals per spectral field
    size_t nphys = 0;   // nxpp*ny*nzp reals per physical field
    size_t plane_s = 0; // nx*ny
    size_t plane_p = 0; // nxpp*ny
    int ldm_z = 0;      // padded leading dimension of the nzp x nzp device operators

    // grid module (common.f90:112-123)
    double xl = 0, dx = 0, dy = 0, alpha = 0, beta = 0, aa = 0, bb = 0;
    std::vector<double> wavx, wavy, wavxx, wavyy, zpts, gf, xpts;
    // mats module (common.f90:189-197), column-major like Fortran
    std::vector<double> dor, d, d2;      // nzp*nzp
    std::vector<double> mdz, md2;        // dfdz / d2fdz2 operators (nzp*nzp)
    std::vector<double> diag, smat, simat;  // nzm, nzm*nzm
    // paral module
    std::vector<double> ubar, ubarp, ubarpp, wbar, wbarpp;
    // lifting functions of solve_nhom_helmholtz (helmholtz.f90:240-307)
    std::vector<double> hbot, htop, d2hbot, d2htop;
    // static Smagorinsky / sigma filter widths (sgsterm.f90:95-103, 158-164, 367-376)
    std::vector<double> delta_static, delta_dyn, delta_sigma;
    // filter_q_3pt z weights for (a,b) and (1,1) (filters.f90:396-448)
    std::vector<double> fz_ab, fz_11;    // 3*nzp each: az, bz, cz
    std::vector<double> fz_16, fz_12;    // (1, 6) and (1, 2): the TNS filters

    // device copies
    int device = 0;
    hipStream_t stream = nullptr;
    double* d_wavx = nullptr;   // nx
    double* d_wavy = nullptr;   // ny
    double* d_zpts = nullptr;   // nzp
    double* d_mdz = nullptr;    // nzp x nzp column-major (ld = nzp)
    double* d_md2 = nullptr;    // nzp x nzp
    double* d_md2b = nullptr;   // rows {0, nzp-1} of md2, 2 x nzp column-major
    double* d_d = nullptr;      // nzp x nzp (matmul path for nz <= 31)
    double* d_smat = nullptr;   // nzm x nzm
    double* d_simat = nullptr;  // nzm x nzm
    double* d_diag = nullptr;   // nzm
    double* d_lift = nullptr;   // 4*nzp: hbot, htop, d2hbot, d2htop; then 4*nzm: Einv of each
    double* d_paral = nullptr;  // 3*nzp: ubar, ubarp, ubarpp
    double* d_delta = nullptr;  // 3*nzp: static, dynamic, sigma
    double* d_fz = nullptr;     // 12*nzp: (a,b), (1,1), (1,6), (1,2) weights
    double* d_smag = nullptr;   // nzp (static Smagorinsky -(cs delta fvd)^2) + scratch
    double* d_red = nullptr;    // reduction scratch (kRedDoubles doubles)
    double* d_ftab = nullptr;   // spectral filter factors of one plane (lh*ny)
    double* d_diag_red = nullptr;  // diagnostics partials + small tables (kDiagDoubles)
    double* d_lam = nullptr;    // sponge lambda(x), nxpp (forcing = 1; pads 0)
    double sponge[3] = {0.15, 0.25, 0.70};  // fsize, lfsize, sfsize (run/input.dat values)
    int* d_walls = nullptr;     // wall planes of the static model's utau average
    int nwalls = 0;
    // parity-split forms of dfdz (d_mdz) and of the two Helmholtz projections, with
    // the eigenvalues and Einv-lifting data in parity-ordered mode rows (d_helm_s:
    // nzm, then 4*nzm); not built (ok = false) on a grid without the symmetry
    lsb::ZSplitOp zs_dz, zs_h1, zs_h2;
    double* d_helm_s = nullptr;
    float* d_mdz32 = nullptr;
    float* d_smat32 = nullptr;
    float* d_simat32 = nullptr;

    lsb::HorPlan hp;            // nx, ny
    lsb::HorPlan hp_lg;         // dealiasing sizes
    int nx_lg = 0, ny_lg = 0;

    // staging buffers for tier A (device)
    std::vector<void*> stage;
    std::vector<size_t> stage_bytes;

    // tier B state
    bool state_ready = false;
    double* st = nullptr;       // contiguous block of all state + work fields
    size_t st_bytes = 0;
    // istep: the run's step counter (main.f90:61 loop variable: start-up coefficients,
    // adaptive-dt schedule); isum: steps taken in total (time module its/isum, the
    // snapshot and restart headers).  coef_valid is cleared when state is loaded.
    int istep = 0, isum = 0;
    bool coef_valid = false;
    double t = 0.0, dt = 0.0, dt2 = 0.0, dt3 = 0.0;
    // adaptive time step (navierstokes.f90:74-87; runparam dtadapt, iomod, fsize)
    int dtadapt = 0, iomod = 1;
    double fsize = 0.0;
    // filtering (navierstokes.f90:31-39; runparam ifilt, modfilt)
    int ifilt = 0, modfilt = 1;
    // on-the-fly time average (main.f90:69-76; runparam otfavg, istartavg, iomod):
    // uf, vf, wf, tempf, pf averages, 5 spectral FP64 fields in one block
    int otfavg = 0, istartavg = 0, avg_iomod = 1;
    double* d_tavg = nullptr;
    double abfac[3] = {0, 0, 0};
    double bdf[3] = {0, 0, 0};
    double bdfv = 0.0;
    // rotating slots (indices into the state block)
    int s_u[3], s_un[3], s_unm[3], s_nl[3], s_nlm[3], s_p, s_pn;
    double* bc_spec = nullptr;  // spectral bound() field, planes 0 and nzp-1 (2*plane_s)

    // member of a captured ensemble (lsb_ensemble_create): its steps run from graphs and
    // the staging buffers the graphs read may not be reallocated
    bool graphs_live = false;
    bool stage_locked = false;
    bool capturing = false;                 // stage slots used while set are the graphs'
    std::vector<char> stage_in_graph;
    // launch-shape variants chosen by lsb_autotune (-1: the built-in default)
    int tune[lsb::TUNE_SLOTS] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    float tune_ms[lsb::TUNE_SLOTS] = {};
    // profiling
    bool profiling = false;
    unsigned prof_mask = ~0u;  // classes bracketed while profiling
    struct ProfEv {
        hipEvent_t a, b;
        int cls;
        double flops, bytes;
    };
    std::vector<ProfEv> prof;
    double prof_ms[4] = {0, 0, 0, 0};
    int prof_n[4] = {0, 0, 0, 0};
};

namespace lsb {

// host setup (setup.cpp)
int build_grid(lsb_ctx* c);
int build_operators(lsb_ctx* c);
int build_derived(lsb_ctx* c);
void cheb_matrix(int nzp, double xl, const double* gf, double* dor, double* d, double* d2);
void dct_diff_matrix(int n, double xl, const double* gf, int p, double* m, int exact = 0);
int eigen_decompose(const double* a, int n, double* lam, double* e, double* ei);
void lifting_functions(lsb_ctx* c);
void sponge_lambda(int nx, const double* xpts, double xlen, double u0, double fsize, double lfsize, double sfsize,
                   double* lam);
int orrsomm_solve(int nzp, const double* d, const double* d2, const double* ubar, const double* ubarp,
                  const double* ubarpp, const double* zpts, double xnu, double eps, double alphax, double betax,
                  double* omega, double* uvec, double* vvec, double* wvec);

// device memory helpers (capi.cpp)
int dev_alloc(void** p, size_t bytes);
int upload(void* dst, const void* src, size_t bytes, hipStream_t s);
int stage_get(lsb_ctx* c, int slot, size_t bytes, void** p);
int upload_context(lsb_ctx* c);
int zsplit_setup(lsb_ctx* c, const std::vector<double>& elift);

// snapshot / restart records and per-step diagnostics (iodiag.hip)
// phys: 5 physical FP64 fields u, v, w, temp, p (null = zero); add_u / add_t: per-plane
// profiles added to u / temp (null = none); out: 5*nx*ny*nzp floats in (x, z, y) order.
int snap_pack(lsb_ctx* c, const double* const* phys, const double* add_u, const double* add_t, float* out,
              hipStream_t s);
// nf spectral FP64 fields (null = zero) <-> nf FP32 blocks in (z, y, x) order, z fastest.
int restart_pack(lsb_ctx* c, const double* const* spec, int nf, float* out, hipStream_t s);
int restart_unpack(lsb_ctx* c, const float* in, double* const* spec, int nf, hipStream_t s);
// extrema / CFL partials of physical u, v, w, temp (temp, tbar may be null); stats[11]:
// umax, umin, vmax, vmin, wmax, wmin, tmax, tmin, max|u+ubar|, max|v|, z-CFL.  Synchronises.
int maxcalc_reduce(lsb_ctx* c, const double* u, const double* v, const double* w, const double* temp,
                   const double* tbar_host, double dt, int findex, double* stats, hipStream_t s);
int nan_scan(lsb_ctx* c, const double* const* f, int nf, size_t n, bool* found, hipStream_t s);
// spanwise means of two physical planes -> host (2*nx doubles).  Synchronises.
int wall_mean(lsb_ctx* c, const double* phys2, double* mean_host, hipStream_t s);
// per-plane sums of 0.5 (u^2 + v^2 + w^2) of physical fields (u + ubar when add_ubar) -> host
int plane_ke(lsb_ctx* c, const double* u, const double* v, const double* w, bool add_ubar, double* ee_host,
             hipStream_t s);
// zero the (kx, ky) = (0, 0) mode of nf spectral fields
int zero_mean_modes(lsb_ctx* c, double* const* spec, int nf, hipStream_t s);
// avg_l += f_l / div for nf spectral FP64 fields (the on-the-fly time average)
int tavg_accum(lsb_ctx* c, const double* const* f, double* const* avg, int nf, int div, hipStream_t s);
// time average after a step when otfavg is on and the total step count exceeds istartavg (capi.cpp)
int state_time_average(lsb_ctx* c);
// NaN check + maxcalc on the resident state after a step when dtadapt is on (capi.cpp)
int state_adapt_dt(lsb_ctx* c);
// filterall3pt / TNS filtering of the resident state before a step when ifilt >= 1 (capi.cpp)
int state_filter_hook(lsb_ctx* c);

// temperature / buoyancy path of the FP64 step (temperature.hip)
bool temp_on(const lsb_ctx* c);
int temp_enable(lsb_ctx* c);
void temp_free(lsb_ctx* c);
int temp_set_params(lsb_ctx* c, const double* tbar, const double* tbarp, const double* tbarpp, double coex,
                    double gx, double gy, double gz);
const double* temp_tbar_dev(const lsb_ctx* c);   // tbar, tbarp, tbarpp (3 nzp) or null
const double* temp_tbar_host(const lsb_ctx* c);  // tbar (nzp) or null
double* temp_field(lsb_ctx* c, int which);       // 0 tempf, 1 tnf, 2 tnmf, 3 nltnf, 4 nltnmf
int temp_nlterm(lsb_ctx* c, const double* const* uif, bool filter_u, hipStream_t s);
int temp_update(lsb_ctx* c, hipStream_t s);
int temp_buoyancy(lsb_ctx* c, double* const* rhs, hipStream_t s);

// profiling helpers
// profiling classes: 2-D FFTs, Chebyshev contractions, Helmholtz projections, the RHS
// assembly, the physical-space products/SGS pass, other pointwise kernels
enum { PC_FFT = 0, PC_Z = 1, PC_HELM = 2, PC_PW = 3, PC_RHS = 4, PC_NL = 5, kProfClasses = 6 };
int prof_mark_begin(lsb_ctx* c, int cls, hipStream_t s, int* token, double flops = 0.0, double bytes = 0.0);
int prof_mark_end(lsb_ctx* c, int token, hipStream_t s);

// FFT (fft.hip)
int hor_plan_create(HorPlan* hp, int nx, int ny);
void hor_plan_destroy(HorPlan* hp);
// nf physical fields (row stride nxpp) -> spectral (row stride nx); batched.
int horfft_fwd(lsb_ctx* c, const HorPlan& hp, const double* const* phys, double* const* spec, int nf,
               int nplanes, hipStream_t s);
int horfft_inv(lsb_ctx* c, const HorPlan& hp, const double* const* spec, double* const* phys, int nf,
               int nplanes, hipStream_t s);
// inverse with per-field d/dx, d/dy applied on load (FOP_*); fast lengths only
bool horfft_has_ops(const HorPlan& hp);
template <class T>
int horfft_xfwd(lsb_ctx* c, const HorPlan& hp, const T* const* phys, T* const* rows, int nf, int nplanes,
                hipStream_t s);
struct YRhsArgs;
template <class R, class TH>
int horfft_yfwd_rhs(lsb_ctx* c, const HorPlan& hp, R* const* pieces, int nplanes, const YRhsArgs& a, hipStream_t s);
// scratch: the y pass's staging buffer (nf * nplanes * ny * nx doubles) supplied by the
// caller; null takes the context's staging slot 8 (allocated on first use)
int horfft_inv_ops(lsb_ctx* c, const HorPlan& hp, const double* const* spec, double* const* phys, int nf,
                   int nplanes, const int* ops, hipStream_t s, double* scratch = nullptr);
// allocation-free entry points (noalloc.hip): resolve the device address of the
// file-scope scratch array on `device` (the current device); called by lsb_create
int noalloc_resolve(int device);
// u, du/dx, du/dy of 3 spectra (one load each); RT = physical storage and arithmetic type
template <class RT>
int horfft_inv_grad(lsb_ctx* c, const HorPlan& hp, const double* const* uif, RT* const* u, RT* const* dudx,
                    RT* const* dudy, int nplanes, const double* ftab, hipStream_t s);
// mixed form: FP32 physical fields, FP64 spectra
template <class TO>
int horfft_fwd_mixed(lsb_ctx* c, const HorPlan& hp, const float* const* phys, TO* const* spec, int nf, int nplanes,
                     hipStream_t s);
template <class TI>
int horfft_inv_mixed(lsb_ctx* c, const HorPlan& hp, const TI* const* spec, float* const* phys, int nf, int nplanes,
                     const int* ops, hipStream_t s);

// filters.f90:281-335 specfilter_exp_xy on one mode: rebuilt from its magnitude and
// phase, fmag*filt*(cos(phase), sin(phase)) with cos(phase) = re/fmag and the phase
// signed by im; cos(acos(c)) = c and |sin(acos(c))| = sqrt((1-c)(1+c)) are the same
// functions of the rounded c, so this keeps the reference's loss of Im parts that are
// tiny next to Re without the transcendental round trip.
__host__ __device__ inline void specfilter_mode(double& re, double& im, double filt) {
    const double fmag = sqrt(re * re + im * im);
    if (fmag == 0.0) {
        re = 0.0;
        im = 0.0;
        return;
    }
    const double cosc = re / fmag;
    double sn = sqrt(fmax((1.0 - cosc) * (1.0 + cosc), 0.0));
    if (im < 0.0) sn = -sn;
    const double fmagf = fmag * filt;
    re = fmagf * cosc;
    im = fmagf * sn;
}

// register-resident power-of-two passes (fft_fast.hip); R arithmetic, TI/TO
// element types of input and output
enum { FOP_NONE = 0, FOP_DX = 1, FOP_DY = 2 };
struct FftOps {
    int op[kMaxFields];
};
bool fast_fft_len(int L);
template <class R, class TI, class TO>
int fast_xfwd(const FieldList<TI>& in, const FieldList<TO>& out, int nrows, int nf, int lh, int istride,
              int ostride, const cpx<R>* tw, const cpx<R>* twr, hipStream_t s);
template <class R, class TI, class TO>
int fast_xinv(const FieldList<TI>& in, const FieldList<TO>& out, int nrows, int nf, int lh, int istride, int nxpp,
              const cpx<R>* tw, const cpx<R>* twr, hipStream_t s);
template <class R, class TI, class TO>
int fast_yfwd(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny,
              const cpx<R>* tw, R scale, hipStream_t s, int cols = 0);
template <class R, class TI, class TO>
int fast_yinv(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny, int ph,
              const cpx<R>* tw, const FftOps& ops, const double* wavx, const double* wavy, hipStream_t s,
              int cols = 0);
// nsrc spectra -> 3 nsrc outputs (u, du/dx, du/dy per source), one load per source
// ftab (lh*ny factors, index m + lh*ky) applies the spectral filter on load; null: none
template <class R, class TI, class TO>
int fast_yinv3(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nsrc, int lh, int ny, int ph,
               const cpx<R>* tw, const double* wavx, const double* wavy, const double* ftab, hipStream_t s,
               int cols = 0);

// Forward y pass of the three pieces of N_l (nlt_l, h_(l,x), h_(l,y), as x-transformed
// rows) fused with the EX3/BDF3 velocity right-hand side: N_l and rhs_l are written,
// the transformed pieces never leave registers.  pieces.p[3 l + q], q = nlt, h_x, h_y.
struct YRhsArgs {
    const double* wavx;    // nx (kx at 2 m)
    const double* wavy;    // ny
    const double* ubarpp;  // nzp
    void* nout[3];         // N_l (TH)
    const void* nl[3];     // N history (TH)
    const void* nlm[3];
    const double* u[3];
    const double* un[3];
    const double* unm[3];
    double* rhs[3];
    double ab0, ab1, ab2, bdf0, bdf1, bdf2, dt, dtnui;
};
template <class R, class TH>
int fast_yfwd_rhs(const FieldList<R>& pieces, int nplanes, int lh, int ny, const cpx<R>* tw, R scale,
                  const YRhsArgs& a, hipStream_t s, int cols = 0);
// times the kx-column tiles of the y passes on this context's grid (scratch `work` of
// at least 24 physical fields of doubles) and records the fastest in the tuning table
int fft_tune_y(lsb_ctx* c, void* work, hipStream_t s);

// z-contractions (zops.hip)
// out_f = M * in_f along z for nf fields; M column-major nout x nin (ld = ldm).
int zcontract(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const double* const* in,
              double* const* out, int nf, hipStream_t s);
// pressure right-hand side -d(div_h)/dz - k^2 w in one contraction
int zcontract_press(lsb_ctx* c, const double* divh, const double* w, double* rhs, hipStream_t s);
// fused pressure-step contractions (parity-split dfdz only; zsplit_dz tells)
bool zsplit_dz(const lsb_ctx* c);
int zcontract_pressdiv(lsb_ctx* c, const double* u, const double* v, const double* w, double* rhs, hipStream_t s);
int zcontract_proj(lsb_ctx* c, const double* phi, double* p, double* u, double* v, double dt, double bdfv,
                   hipStream_t s);
// the whole pressure step in one kernel; LSB_E_UNSUPPORTED when it does not apply
int zpress_step(lsb_ctx* c, double* u, double* v, const double* w, double* wout, double* p, const double* bc_top,
                const double* bc_bot, double dt, double bdfv, hipStream_t s);
// mixed form: FP32 matrix-core contraction, panel TI, output TO (FP32 products and sums)
// P columns per plane (0: spectral nx*ny); with aux, out = aux - M in (out may alias aux)
template <class TI, class TO>
int zcontract_f32(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const TI* const* in, TO* const* out,
                  int nf, hipStream_t s, long P = 0, const TO* const* aux = nullptr);
int zcontract_ex(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const double* const* in,
                 double* const* out, int nf, long P, const double* const* aux, hipStream_t s);
// Helmholtz solves in place: rhs_f (nzp planes); bc_f may be null (homogeneous).
// sigma = dtnui (constant over fields).
// split32: the mixed form's velocity solves (FP64 leading term, FP32 correction) when
// sigma dominates the eigenvalues; otherwise the FP64 projections
int helmholtz(lsb_ctx* c, double* const* rhs, const double* const* bc_top,
              const double* const* bc_bot, int nf, double sigma, hipStream_t s, bool split32 = false);

// spectral pointwise (spectral.hip)
int sp_dfdx(lsb_ctx* c, const double* in, double* out, hipStream_t s);
int sp_dfdy(lsb_ctx* c, const double* in, double* out, hipStream_t s);
int sp_divuv(lsb_ctx* c, const double* u, const double* v, double* out, hipStream_t s);
int sp_norm(lsb_ctx* c, double* const* f, int nf, hipStream_t s);
int sp_lapl_combine(lsb_ctx* c, const double* d2u, const double* u, double* out, hipStream_t s);
int sp_axpby(lsb_ctx* c, double* y, const double* x, double a, double b, hipStream_t s);

// physical (physical.hip)
int ph_filter3(lsb_ctx* c, const double* const* in, double* const* out, int nf, double a, double b,
               int interpflag, double* tmp, hipStream_t s);
int ph_sgs(lsb_ctx* c, int model, const double* ui, const double* uiuj, const double* g, double* tau,
           hipStream_t s);

// step (step.hip)
int step_alloc(lsb_ctx* c);
int step_run(lsb_ctx* c, int nsteps, bool with_norm);
// graph capture of the step (ensemble.hip): one step enqueued with its host bookkeeping,
// the bookkeeping alone, the slot-rotation period, and whether a step can be captured
int step_enqueue(lsb_ctx* c);
void step_host_advance(lsb_ctx* c);
int step_rotation_period(lsb_ctx* c);
bool step_capturable(const lsb_ctx* c, std::string& why);
int step_prepare(lsb_ctx* c);
void step_free(lsb_ctx* c);
double* step_field(lsb_ctx* c, int which);
bool step_hist32(const lsb_ctx* c, int which);
int step_hist_set(lsb_ctx* c, int which, const double* src, hipStream_t s);
int step_hist_get(lsb_ctx* c, int which, double* dst, hipStream_t s);
int nlterm_dev(lsb_ctx* c, const double* const* uif, double* const* nlt, hipStream_t s);
// full 9-component tensors (tier A): work must hold 69 (dynamic) / 40 (interscale) physical fields
int dyn_full(lsb_ctx* c, double* ui, double* uiuj, double* g, double* tau, double* work, hipStream_t s);
int isc_full(lsb_ctx* c, double* ui, double* tau, double* work, hipStream_t s);

// dealiasing (dealias.hip)
int dealias_setup(lsb_ctx* c);
int dealias_product(lsb_ctx* c, const double* a, const double* b, double* ab, hipStream_t s);
// the same product of two spectral fields, spectral in and out (abf may alias af or bf)
int dealias_product_spec(lsb_ctx* c, const double* af, const double* bf, double* abf, hipStream_t s);
// zinteg on device: n1 x n2 pencils of nzp values (Fortran (n1, n2, nzp)); w: nzp
// per-point weights (per = 1) or null (trapezoid on zpts)
int zinteg_dev(lsb_ctx* c, const double* tmp, double* tint, int n1, int n2, int mapping, int per, int rev,
               const double* w, hipStream_t s);

}  // namespace lsb
