// Pointwise kernels on spectral fields (one thread per complex (kx,ky,z) mode).
//
// Spectral layout: [nzp][ny][nx/2] complex; pair index pr -> m = pr % lh (kx),
// j = (pr / lh) % ny (ky), k = pr / (lh*ny) (z).  wavx is stored per real
// index like the reference (wavx(2m-1) = wavx(2m)), so kx(m) = wavx[2m].

#include <cmath>

#include "lsb_internal.h"
#include "spectral.h"

namespace lsb {

struct SpDims {
    long npairs;  // nzp*ny*lh
    int lh, ny;
    long plane_pairs;
};

static SpDims sp_dims(const lsb_ctx* c, int nplanes) {
    SpDims d;
    d.lh = c->nx / 2;
    d.ny = c->ny;
    d.plane_pairs = (long)d.lh * d.ny;
    d.npairs = d.plane_pairs * nplanes;
    return d;
}

static dim3 sp_grid(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 8) b = 65535L * 8;
    return dim3((unsigned)b);
}

#define SP_LOOP(n) for (long pr = (long)blockIdx.x * blockDim.x + threadIdx.x; pr < (n); pr += (long)gridDim.x * blockDim.x)

typedef double2 c2;

// derivatives.f90:261-303: d/dx -> (-kx im, kx re), d/dy -> (-ky im, ky re)
__global__ void k_dfdx(SpDims d, const double* __restrict__ wavx, const c2* __restrict__ in, c2* __restrict__ out) {
    SP_LOOP(d.npairs) {
        const int m = (int)(pr % d.lh);
        const double w = wavx[2 * m];
        const c2 u = in[pr];
        out[pr] = make_double2(-w * u.y, w * u.x);
    }
}
__global__ void k_dfdy(SpDims d, const double* __restrict__ wavy, const c2* __restrict__ in, c2* __restrict__ out) {
    SP_LOOP(d.npairs) {
        const int j = (int)((pr / d.lh) % d.ny);
        const double w = wavy[j];
        const c2 u = in[pr];
        out[pr] = make_double2(-w * u.y, w * u.x);
    }
}
// derivatives.f90:394-409
__global__ void k_divuv(SpDims d, const double* __restrict__ wavx, const double* __restrict__ wavy,
                        const c2* __restrict__ u, const c2* __restrict__ v, c2* __restrict__ out) {
    SP_LOOP(d.npairs) {
        const int m = (int)(pr % d.lh);
        const int j = (int)((pr / d.lh) % d.ny);
        const double wx = wavx[2 * m], wy = wavy[j];
        const c2 a = u[pr], b = v[pr];
        out[pr] = make_double2(-wx * a.y - wy * b.y, wx * a.x + wy * b.x);
    }
}
// derivatives.f90:306-326 lapl: d2u - (kx^2 + ky^2) u
__global__ void k_lapl(SpDims d, const double* __restrict__ wavx, const double* __restrict__ wavy,
                       const c2* __restrict__ d2u, const c2* __restrict__ u, c2* __restrict__ out) {
    SP_LOOP(d.npairs) {
        const int m = (int)(pr % d.lh);
        const int j = (int)((pr / d.lh) % d.ny);
        const double wx = wavx[2 * m], wy = wavy[j];
        const double w2 = wx * wx + wy * wy;
        const c2 a = d2u[pr], b = u[pr];
        out[pr] = make_double2(a.x - w2 * b.x, a.y - w2 * b.y);
    }
}

// filters.f90:338-360 specfilter_fourier2d: exp(-amp theta^(2p)) for |k| >= kc*kmax,
// tabulated once per call for the (kx, ky) modes of one plane.
__global__ void k_filt_table(SpDims d, const double* __restrict__ wavx, const double* __restrict__ wavy,
                             double* __restrict__ tab, double akmax, double kc, double p, double amp, int filtery) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int)d.plane_pairs) return;
    const int m = t % d.lh, j = t / d.lh;
    const double wx = wavx[2 * m], wy = wavy[j];
    const double ak = filtery == 1 ? sqrt(wx * wx + wy * wy) : wx;
    const double akc = akmax * kc;
    double filt = 1.0;
    if (fabs(ak) >= akc) {
        const double theta = (fabs(ak) - akc) / (akmax - akc);
        filt = exp(-amp * pow(theta, 2.0 * p));
    }
    tab[t] = filt;
}

// filters.f90:281-335 specfilter_exp_xy (specfilter_mode per element).  One plane
// per blockIdx.y, no 64-bit index division.
__global__ void k_specfilter(SpDims d, const double* __restrict__ tab, FieldList<double> in, FieldList<double> out,
                             int nzp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int)d.plane_pairs) return;
    const int f = blockIdx.y / nzp;
    const long q = (long)(blockIdx.y - f * nzp) * d.plane_pairs + t;
    const c2 u = reinterpret_cast<const c2*>(in.p[f])[q];
    double re = u.x, im = u.y;
    specfilter_mode(re, im, tab[t]);
    reinterpret_cast<c2*>(out.p[f])[q] = make_double2(re, im);
}

// horfft.f90:153-224 norm on the kx = 0 column and the ky-Nyquist row.
// One thread per (field, z, j) with j in 0..ny-1 for the Nyquist row and the
// symmetric pairs handled by the thread owning the lower index.
__global__ void k_norm(FieldList<double> f, int nf, int nzp, int ny, int nx, int kx0_zero) {
    const int nyh = ny / 2;
    const long n = (long)nf * nzp * ny;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t % ny);
        const long kf = t / ny;
        const int k = (int)(kf % nzp);
        const int fi = (int)(kf / nzp);
        double* pl = f.p[fi] + (long)k * nx * ny;
        if (kx0_zero) {
            if (j == 0) pl[1] = 0.0;
            if (ny > 2 && j >= 1 && j <= nyh - 1) {
                const int jp = ny - j;  // 0-based partner of j (ny + 2 - (j+1) - 1)
                double* a = pl + (long)j * nx;
                double* b = pl + (long)jp * nx;
                const double re = 0.5 * (a[0] + b[0]);
                const double im = 0.5 * (a[1] - b[1]);
                a[0] = re;
                a[1] = im;
                b[0] = re;
                b[1] = -im;
            }
        }
    }
}
__global__ void k_norm_nyq(FieldList<double> f, int nf, int nzp, int ny, int nx) {
    const int nyh = ny / 2;
    const long n = (long)nf * nzp * nx;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const long kf = t / nx;
        const int k = (int)(kf % nzp);
        const int fi = (int)(kf / nzp);
        f.p[fi][((long)k * ny + nyh) * nx + i] = 0.0;
    }
}

// ---------------------------------------------------------------- launchers
int sp_dfdx(lsb_ctx* c, const double* in, double* out, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    hipLaunchKernelGGL(k_dfdx, sp_grid(d.npairs), dim3(256), 0, s, d, c->d_wavx, (const c2*)in, (c2*)out);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int sp_dfdy(lsb_ctx* c, const double* in, double* out, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    hipLaunchKernelGGL(k_dfdy, sp_grid(d.npairs), dim3(256), 0, s, d, c->d_wavy, (const c2*)in, (c2*)out);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int sp_divuv(lsb_ctx* c, const double* u, const double* v, double* out, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    int tok = -1;
    prof_mark_begin(c, PC_PW, s, &tok, 0.0, 3.0 * 16.0 * (double)d.npairs);
    hipLaunchKernelGGL(k_divuv, sp_grid(d.npairs), dim3(256), 0, s, d, c->d_wavx, c->d_wavy, (const c2*)u,
                       (const c2*)v, (c2*)out);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int sp_lapl_combine(lsb_ctx* c, const double* d2u, const double* u, double* out, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    hipLaunchKernelGGL(k_lapl, sp_grid(d.npairs), dim3(256), 0, s, d, c->d_wavx, c->d_wavy, (const c2*)d2u,
                       (const c2*)u, (c2*)out);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
// Factors exp(-amp theta^(2p)) of one plane into c->d_ftab (returned).
const double* sp_filter_table(lsb_ctx* c, double kc, double p, double amp, int filtery, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    const double wavxmax = c->nxh * c->alpha;
    const double wavymax = c->nyh * c->beta;
    double akmax = wavxmax;
    if (filtery == 1) akmax = sqrt(wavxmax * wavxmax + wavymax * wavymax);
    if (!c->d_ftab && dev_alloc((void**)&c->d_ftab, (size_t)d.plane_pairs * sizeof(double)) != LSB_OK) return nullptr;
    const unsigned nb = (unsigned)((d.plane_pairs + 255) / 256);
    hipLaunchKernelGGL(k_filt_table, dim3(nb), dim3(256), 0, s, d, c->d_wavx, c->d_wavy, c->d_ftab, akmax, kc, p, amp,
                       filtery);
    if (hipGetLastError() != hipSuccess) return nullptr;
    return c->d_ftab;
}

int sp_specfilter(lsb_ctx* c, const double* const* in, double* const* out, int nf, double kc, double p, double amp,
                  int filtery, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    FieldList<double> fi{}, fo{};
    for (int f = 0; f < nf; ++f) {
        fi.p[f] = const_cast<double*>(in[f]);
        fo.p[f] = out[f];
    }
    int tok = -1;
    prof_mark_begin(c, PC_PW, s, &tok, 0.0, 2.0 * (double)d.npairs * nf * 2 * sizeof(double));
    const double* tab = sp_filter_table(c, kc, p, amp, filtery, s);
    if (!tab) {
        set_error("specfilter: table allocation/launch failed");
        return LSB_E_HIP;
    }
    const unsigned nb = (unsigned)((d.plane_pairs + 255) / 256);
    hipLaunchKernelGGL(k_specfilter, dim3(nb, (unsigned)(c->nzp * nf)), dim3(256), 0, s, d, tab, fi, fo, c->nzp);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int sp_norm(lsb_ctx* c, double* const* f, int nf, hipStream_t s) {
    FieldList<double> fl{};
    for (int i = 0; i < nf; ++i) fl.p[i] = f[i];
    const long n1 = (long)nf * c->nzp * c->ny;
    const int kx0 = (c->wavx[0] == 0.0) ? 1 : 0;
    hipLaunchKernelGGL(k_norm, sp_grid(n1), dim3(256), 0, s, fl, nf, c->nzp, c->ny, c->nx, kx0);
    LSB_HIP(hipGetLastError());
    if (c->ny > 2) {
        const long n2 = (long)nf * c->nzp * c->nx;
        hipLaunchKernelGGL(k_norm_nyq, sp_grid(n2), dim3(256), 0, s, fl, nf, c->nzp, c->ny, c->nx);
        LSB_HIP(hipGetLastError());
    }
    return LSB_OK;
}

// ------------------------------------------------------- step kernels
// N_l = nlt_l - dfdx(h_l0) - dfdy(h_l1) - dhdz_l   (navierstokes.f90:781-795),
// then EX3/BDF3 right-hand side (navierstokes.f90:1236-1296):
//   rhs_l = -dtnui*( dt*(ab0 N + ab1 Nn + ab2 Nnm) + bdf0 u + bdf1 un + bdf2 unm )
//   rhs_0(kx=0,ky=0,z) -= ubarpp(z)
// N_l is also written (it becomes the next step's nlnif).
template <class TN>
__device__ __forceinline__ c2 ld_n(const void* p, long q) {
    const TN* t = reinterpret_cast<const TN*>(p) + 2 * q;
    return make_double2((double)t[0], (double)t[1]);
}

template <class TN>
__device__ __forceinline__ void st_n(void* p, long q, double re, double im) {
    TN* t = reinterpret_cast<TN*>(p) + 2 * q;
    t[0] = (TN)re;
    t[1] = (TN)im;
}

template <class TN, class TH>
__global__ void k_rhs_vel(SpDims d, const double* __restrict__ wavx, const double* __restrict__ wavy, StepRhsArgs a) {
    const long n = d.npairs * 3;
    SP_LOOP(n) {
        const int l = (int)(pr / d.npairs);
        const long q = pr - (long)l * d.npairs;
        const int m = (int)(q % d.lh);
        const long jk = q / d.lh;
        const int j = (int)(jk % d.ny);
        const int k = (int)(jk / d.ny);
        const double wx = wavx[2 * m], wy = wavy[j];
        const c2 nl = ld_n<TN>(a.nlt[l], q);
        const c2 h0 = ld_n<TN>(a.h0[l], q);
        const c2 h1 = ld_n<TN>(a.h1[l], q);
        const c2 hz = a.hz[l] ? ld_n<TN>(a.hz[l], q) : make_double2(0.0, 0.0);
        double nr = nl.x, ni = nl.y;
        nr = nr - (-wx * h0.y);
        ni = ni - (wx * h0.x);
        nr = nr - (-wy * h1.y);
        ni = ni - (wy * h1.x);
        nr = nr - hz.x;
        ni = ni - hz.y;
        if (sizeof(TH) < sizeof(double)) {
            // the FP32 history holds N rounded once; the right-hand side uses the same
            // rounded value this step as the next two do
            nr = (double)(TH)nr;
            ni = (double)(TH)ni;
        }
        st_n<TH>(a.nout[l], q, nr, ni);
        const c2 n1 = ld_n<TH>(a.nl[l], q);
        const c2 n2 = ld_n<TH>(a.nlm[l], q);
        const c2 u0 = reinterpret_cast<const c2*>(a.u[l])[q];
        const c2 u1 = reinterpret_cast<const c2*>(a.un[l])[q];
        const c2 u2 = reinterpret_cast<const c2*>(a.unm[l])[q];
        const double cr = a.ab0 * nr + a.ab1 * n1.x + a.ab2 * n2.x;
        const double ci = a.ab0 * ni + a.ab1 * n1.y + a.ab2 * n2.y;
        double rr = -a.dtnui * (a.dt * cr + a.bdf0 * u0.x + a.bdf1 * u1.x + a.bdf2 * u2.x);
        double ri = -a.dtnui * (a.dt * ci + a.bdf0 * u0.y + a.bdf1 * u1.y + a.bdf2 * u2.y);
        if (l == 0 && m == 0 && j == 0) rr = rr - a.ubarpp[k];
        reinterpret_cast<c2*>(a.rhs[l])[q] = make_double2(rr, ri);
    }
}

// Boundary planes of the viscous step (navierstokes.f90:1321-1373), one thread
// per (kx,ky) mode.  Writes bc_top[l], bc_bot[l] (one plane each).
__global__ void k_bc_vel(SpDims d, const double* __restrict__ wavx, const double* __restrict__ wavy, StepBcArgs a) {
    SP_LOOP(d.plane_pairs) {
        const int m = (int)(pr % d.lh);
        const int j = (int)(pr / d.lh);
        const double wx = wavx[2 * m], wy = wavy[j];
        const double wav2 = wx * wx + wy * wy;
        const double den = wav2 + a.dtnui / a.bdfv;
        const double dtb = a.dt * a.bdfv;
        // top (plane 0): -rhs/(wav2 + dtnui/bdfv) for u and v (free flow), or the wall's
        // Zang-Hussaini values (Dirichlet ceiling), or the free values plus the difference
        // that the ceiling weight blends in physical space (top_bc 2, finished by the caller)
        const c2 r0 = reinterpret_cast<const c2*>(a.rhs0)[pr];
        const c2 r1 = reinterpret_cast<const c2*>(a.rhs1)[pr];
        c2 fu = make_double2(-r0.x / den, -r0.y / den);
        c2 fv = make_double2(-r1.x / den, -r1.y / den);
        if (a.top_bc) {
            const c2 q0 = reinterpret_cast<const c2*>(a.p_top)[pr];
            const c2 q1 = reinterpret_cast<const c2*>(a.pn_top)[pr];
            const double qr = a.pfac0 * q0.x + a.pfac1 * q1.x;
            const double qi = a.pfac0 * q0.y + a.pfac1 * q1.y;
            const c2 du = make_double2(dtb * (-wx * qi), dtb * (wx * qr));
            const c2 dv = make_double2(dtb * (-wy * qi), dtb * (wy * qr));
            if (a.top_bc == 1) {
                fu = du;
                fv = dv;
            } else {
                reinterpret_cast<c2*>(a.cdiff[0])[pr] = make_double2(du.x - fu.x, du.y - fu.y);
                reinterpret_cast<c2*>(a.cdiff[1])[pr] = make_double2(dv.x - fv.x, dv.y - fv.y);
            }
        }
        reinterpret_cast<c2*>(a.bct[0])[pr] = fu;
        reinterpret_cast<c2*>(a.bct[1])[pr] = fv;
        // wall (plane nzp-1): dt*bdfv*d(p*)/dx, dt*bdfv*d(p*)/dy, p* = pfac0 p + pfac1 pn
        const c2 p0 = reinterpret_cast<const c2*>(a.p_wall)[pr];
        const c2 p1 = reinterpret_cast<const c2*>(a.pn_wall)[pr];
        const double psr = a.pfac0 * p0.x + a.pfac1 * p1.x;
        const double psi = a.pfac0 * p0.y + a.pfac1 * p1.y;
        reinterpret_cast<c2*>(a.bcb[0])[pr] = make_double2(dtb * (-wx * psi), dtb * (wx * psr));
        reinterpret_cast<c2*>(a.bcb[1])[pr] = make_double2(dtb * (-wy * psi), dtb * (wy * psr));
        // w: imposed field (bound) plus the Zang-Hussaini correction xnu*dt*bdfv*d2w/dz2
        const c2 d2t = reinterpret_cast<const c2*>(a.d2w)[pr];                // row 0
        const c2 d2b = reinterpret_cast<const c2*>(a.d2w)[d.plane_pairs + pr];  // row nzp-1
        const double zh = a.xnu * a.dt * a.bdfv;
        if (a.nhom) {
            const c2 gt = a.bnd_top ? reinterpret_cast<const c2*>(a.bnd_top)[pr] : make_double2(0.0, 0.0);
            const c2 gb = a.bnd_bot ? reinterpret_cast<const c2*>(a.bnd_bot)[pr] : make_double2(0.0, 0.0);
            reinterpret_cast<c2*>(a.bct[2])[pr] = make_double2(gt.x + zh * d2t.x, gt.y + zh * d2t.y);
            reinterpret_cast<c2*>(a.bcb[2])[pr] = gb;
        } else {
            reinterpret_cast<c2*>(a.bct[2])[pr] = make_double2(zh * d2t.x, zh * d2t.y);
            reinterpret_cast<c2*>(a.bcb[2])[pr] = make_double2(zh * d2b.x, zh * d2b.y);
        }
    }
}

// Projection (navierstokes.f90:1466-1488):
//   p = -(dwdz + divf)/(wav2 dt bdfv) (0 where wav2 = 0); u -= dt bdfv dp/dx; v -= dt bdfv dp/dy
__global__ void k_project(SpDims d, const double* __restrict__ wavx, const double* __restrict__ wavy,
                          const c2* __restrict__ dwdz, const c2* __restrict__ divf, c2* __restrict__ p,
                          c2* __restrict__ u, c2* __restrict__ v, double dt, double bdfv) {
    SP_LOOP(d.npairs) {
        const int m = (int)(pr % d.lh);
        const int j = (int)((pr / d.lh) % d.ny);
        const double wx = wavx[2 * m], wy = wavy[j];
        const double wav2 = wx * wx + wy * wy;
        double pr_ = 0.0, pi_ = 0.0;
        if (wav2 != 0.0) {
            const c2 a = dwdz[pr], b = divf[pr];
            const double den = wav2 * dt * bdfv;
            pr_ = -(a.x + b.x) / den;
            pi_ = -(a.y + b.y) / den;
        }
        p[pr] = make_double2(pr_, pi_);
        const double dtb = dt * bdfv;
        const c2 uu = u[pr], vv = v[pr];
        u[pr] = make_double2(uu.x - dtb * (-wx * pi_), uu.y - dtb * (wx * pr_));
        v[pr] = make_double2(vv.x - dtb * (-wy * pi_), vv.y - dtb * (wy * pr_));
    }
}

int sp_rhs_vel(lsb_ctx* c, const StepRhsArgs& a, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    int tok = -1;
    {   // per component: the N pieces (nlt, h0, h1, hz if present) and two N history
        // fields, three FP64 velocity fields read, N and rhs written
        const double tn = a.n32 ? 4.0 : 8.0, th = a.h32 ? 4.0 : 8.0;
        const double pieces = a.hz[0] ? 4.0 : 3.0;
        const double per_pair = 2.0 * (pieces * tn + 3.0 * th + 3.0 * 8.0 + 8.0);
        prof_mark_begin(c, PC_RHS, s, &tok, 0.0, per_pair * (double)d.npairs * 3.0);
    }
    if (a.n32 && a.h32)
        hipLaunchKernelGGL((k_rhs_vel<float, float>), sp_grid(d.npairs * 3), dim3(256), 0, s, d, c->d_wavx, c->d_wavy,
                           a);
    else if (a.n32)
        hipLaunchKernelGGL((k_rhs_vel<float, double>), sp_grid(d.npairs * 3), dim3(256), 0, s, d, c->d_wavx,
                           c->d_wavy, a);
    else if (!a.h32)
        hipLaunchKernelGGL((k_rhs_vel<double, double>), sp_grid(d.npairs * 3), dim3(256), 0, s, d, c->d_wavx,
                           c->d_wavy, a);
    else  // FP64 pieces into the FP32 history (the mixed form with an FP64 nonlinear chain)
        hipLaunchKernelGGL((k_rhs_vel<double, float>), sp_grid(d.npairs * 3), dim3(256), 0, s, d, c->d_wavx,
                           c->d_wavy, a);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int sp_bc_vel(lsb_ctx* c, const StepBcArgs& a, hipStream_t s) {
    SpDims d = sp_dims(c, 1);
    hipLaunchKernelGGL(k_bc_vel, sp_grid(d.plane_pairs), dim3(256), 0, s, d, c->d_wavx, c->d_wavy, a);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
__global__ void k_row_weight(double* __restrict__ a, double* __restrict__ b, const double* __restrict__ m, int nx,
                             int nxpp, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % nxpp);
    const double w = (i < nx) ? m[i] : 0.0;
    a[t] *= w;
    b[t] *= w;
}

__global__ void k_add_planes(double* __restrict__ y0, double* __restrict__ y1, const double* __restrict__ x0,
                             const double* __restrict__ x1, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y0[t] += x0[t];
    y1[t] += x1[t];
}

int sp_ceiling_blend(lsb_ctx* c, double* const* bct, double* const* cdiff, double* const* phys, const double* m,
                     hipStream_t s) {
    const double* in[2] = {cdiff[0], cdiff[1]};
    LSB_TRY(horfft_inv(c, c->hp, in, phys, 2, 1, s));
    const long np = (long)c->plane_p;
    hipLaunchKernelGGL(k_row_weight, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, phys[0], phys[1], m, c->nx,
                       c->nxpp, np);
    LSB_HIP(hipGetLastError());
    const double* pin[2] = {phys[0], phys[1]};
    LSB_TRY(horfft_fwd(c, c->hp, pin, cdiff, 2, 1, s));
    const long ns = (long)c->plane_s;
    hipLaunchKernelGGL(k_add_planes, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, bct[0], bct[1], cdiff[0],
                       cdiff[1], ns);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int sp_project(lsb_ctx* c, const double* dwdz, const double* divf, double* p, double* u, double* v, double dt,
               double bdfv, hipStream_t s) {
    SpDims d = sp_dims(c, c->nzp);
    int tok = -1;
    // reads dwdz, divf, u, v; writes p, u, v
    prof_mark_begin(c, PC_PW, s, &tok, 0.0, 7.0 * 16.0 * (double)d.npairs);
    hipLaunchKernelGGL(k_project, sp_grid(d.npairs), dim3(256), 0, s, d, c->d_wavx, c->d_wavy, (const c2*)dwdz,
                       (const c2*)divf, (c2*)p, (c2*)u, (c2*)v, dt, bdfv);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

}  // namespace lsb
