// Temperature / buoyancy path of the time step, in FP64 in both forms (the mixed form runs
// it from the FP64 velocity spectra beside its FP32 momentum chain):
//   nlterm_skewsym's temperature part (navierstokes.f90:800-845): grad T of the filtered
//   temperature, products with the full velocity, -1/2 u.grad T - 1/2 div(u T);
//   the temperature update (navierstokes.f90:1251-1275): EX3/BDF3 right-hand side with
//   the base profile's tbarpp, homogeneous Helmholtz solve, history rotation;
//   buoyancy in the velocity right-hand side (navierstokes.f90:1298-1303).
// The two "any |T| > 1e-10" tests of the reference (the spectral filter of T at
// navierstokes.f90:1215 and the buoyancy) are evaluated into device flags, so the step
// stays free of host synchronisation.

#include <cmath>

#include "lsb_internal.h"
#include "spectral.h"

namespace lsb {

namespace {

struct TDims {
    long npairs, plane_pairs, npts;
    int lh, nx, nxpp, ny, nzp;
};

TDims tdims(const lsb_ctx* c) {
    TDims d;
    d.lh = c->nx / 2;
    d.nx = c->nx;
    d.nxpp = c->nxpp;
    d.ny = c->ny;
    d.nzp = c->nzp;
    d.plane_pairs = (long)d.lh * d.ny;
    d.npairs = d.plane_pairs * d.nzp;
    d.npts = (long)d.nx * d.ny * d.nzp;
    return d;
}

dim3 tgrid(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 16) b = 65535L * 16;
    return dim3((unsigned)(b < 1 ? 1 : b));
}

#define T_LOOP(n) for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (n); e += (long)gridDim.x * blockDim.x)

// flag[0] = 1 when some |f| > thr (flag cleared by the caller's fill kernel).  A fixed grid
// of kFlagBlocks blocks of 256 threads, one atomic per block that saw a hit.
constexpr int kFlagBlocks = 1024;
__global__ void __launch_bounds__(256) k_flag_any_gt(const double* __restrict__ f, long n, double thr,
                                                     int* __restrict__ flag) {
    int hit = 0;
    T_LOOP(n) hit |= (fabs(f[e]) > thr);
    __shared__ int sh[4];
    const int any = __any(hit);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = any;
    __syncthreads();
    if (threadIdx.x == 0 && (sh[0] | sh[1] | sh[2] | sh[3])) atomicOr(flag, 1);
}

__global__ void k_fill_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// out = flag ? a : b (spectral, element-wise)
__global__ void k_select(const int* __restrict__ flag, const double* __restrict__ a, const double* __restrict__ b,
                         double* __restrict__ out, long n) {
    const bool f = flag[0] != 0;
    T_LOOP(n) out[e] = f ? a[e] : b[e];
}

// navierstokes.f90:813-830 on the physical points: T + tbar(k); dT/dz + sum(tbarp) (the
// reference adds every tbarp(k) to every plane); nltemp = -1/2 sum_l u_l dT/dx_l with
// u = u' + ubar (l = 0); u_l T.
__global__ void k_temp_products(TDims d, const double* __restrict__ u0, const double* __restrict__ u1,
                                const double* __restrict__ u2, const double* __restrict__ ubar,
                                const double* __restrict__ T, const double* __restrict__ dTx,
                                const double* __restrict__ dTy, const double* __restrict__ dTz,
                                const double* __restrict__ tbar, double sum_tbarp, double* __restrict__ nlt,
                                double* __restrict__ uT0, double* __restrict__ uT1, double* __restrict__ uT2) {
    T_LOOP(d.npts) {
        const int i = (int)(e % d.nx);
        const long jk = e / d.nx;
        const int k = (int)(jk / d.ny);
        const long off = jk * d.nxpp + i;
        const double ua = u0[off] + ubar[k], ub = u1[off], uc = u2[off];
        const double tf = T[off] + tbar[k];
        const double dz = dTz[off] + sum_tbarp;
        double s = 0.0;
        s = s - 0.5 * ua * dTx[off];
        s = s - 0.5 * ub * dTy[off];
        s = s - 0.5 * uc * dz;
        nlt[off] = s;
        uT0[off] = ua * tf;
        uT1[off] = ub * tf;
        uT2[off] = uc * tf;
    }
}

// navierstokes.f90:834-845: nltempf = nltempf - 1/2 d(uT)/dx - 1/2 d(vT)/dy - 1/2 d(wT)/dz
__global__ void k_temp_div(TDims d, const double* __restrict__ wavx, const double* __restrict__ wavy,
                           double2* __restrict__ nlt, const double2* __restrict__ hx, const double2* __restrict__ hy,
                           const double2* __restrict__ dzhz) {
    T_LOOP(d.npairs) {
        const int m = (int)(e % d.lh);
        const int j = (int)((e / d.lh) % d.ny);
        const double wx = wavx[2 * m], wy = wavy[j];
        double2 r = nlt[e];
        const double2 a = hx[e], b = hy[e], cz = dzhz[e];
        // dfdx: (re, im) -> (-kx im, kx re)
        r.x = r.x - 0.5 * (-wx * a.y);
        r.y = r.y - 0.5 * (wx * a.x);
        r.x = r.x - 0.5 * (-wy * b.y);
        r.y = r.y - 0.5 * (wy * b.x);
        r.x = r.x - 0.5 * cz.x;
        r.y = r.y - 0.5 * cz.y;
        nlt[e] = r;
    }
}

// navierstokes.f90:1256-1266: rhs = -dtkapi (dt (ab0 N + ab1 Nn + ab2 Nnm) + bdf0 T + bdf1 Tn
// + bdf2 Tnm) - horfft(tbarpp) (the plane-constant tbarpp is the (0, 0) mode's real part)
__global__ void k_temp_rhs(TDims d, const double* __restrict__ N, const double* __restrict__ Nn,
                           const double* __restrict__ Nnm, const double* __restrict__ T,
                           const double* __restrict__ Tn, const double* __restrict__ Tnm,
                           const double* __restrict__ tbarpp, double ab0, double ab1, double ab2, double bdf0,
                           double bdf1, double bdf2, double dt, double dtkapi, double* __restrict__ rhs) {
    const long n = 2 * d.npairs;
    const long plane = 2 * d.plane_pairs;
    T_LOOP(n) {
        const double nl = ab0 * N[e] + ab1 * Nn[e] + ab2 * Nnm[e];
        double r = -dtkapi * (dt * nl + bdf0 * T[e] + bdf1 * Tn[e] + bdf2 * Tnm[e]);
        if (tbarpp && e % plane == 0) r = r - tbarpp[e / plane];
        rhs[e] = r;
    }
}

// navierstokes.f90:1298-1303: rhs_l -= coex g_l T / xnu when some |T| > 1e-10
__global__ void k_buoyancy(const int* __restrict__ flag, const double* __restrict__ T, double* __restrict__ r0,
                           double* __restrict__ r1, double* __restrict__ r2, double cg0, double cg1, double cg2,
                           double xnu, long n) {
    if (flag[0] == 0) return;
    T_LOOP(n) {
        const double t = T[e];
        r0[e] = r0[e] - cg0 * t / xnu;
        r1[e] = r1[e] - cg1 * t / xnu;
        r2[e] = r2[e] - cg2 * t / xnu;
    }
}

}  // namespace

// ------------------------------------------------------------------ state
// slots of the temperature block (spectral fields first, then physical)
enum {
    TS_T0 = 0,      // 5 rotating state slots: T, Tn, Tnm, Nn, Nnm (index in TempState::s*)
    TS_WORK = 5,    // rhs buffer (rotates into the state)
    TS_NLT = 6,     // nonlinear term of this step (rotates into Nn)
    TS_TF = 7,      // filtered T
    TS_TFILT = 8,   // filter output before the select
    TS_D = 9,       // dT/dx, dT/dy, dT/dz (3)
    TS_H = 12,      // F(uT), F(vT), F(wT) (3)
    TS_DZH = 15,    // d F(wT)/dz
    TS_UF = 16,     // filtered velocity (3)
    TS_NSPEC = 19,
    TP_T = 0,       // physical: T, dTx, dTy, dTz, nltemp, uT, vT, wT, u, v, w
    TP_NPHYS = 11
};

struct TempState {
    double* block = nullptr;
    double* spec[TS_NSPEC];
    double* phys[TP_NPHYS];
    int* flags = nullptr;   // [0] filter test, [1] buoyancy test
    double* d_tbar = nullptr;  // tbar, tbarp, tbarpp (3 nzp)
    std::vector<double> tbar, tbarp, tbarpp;
    double coex = 0.0, g[3] = {0.0, 0.0, 0.0};
    // rotating indices into spec[]
    int T = 0, Tn = 1, Tnm = 2, Nn = 3, Nnm = 4, W = TS_WORK, NL = TS_NLT;
};

static std::vector<std::pair<const lsb_ctx*, TempState*>> g_temp;

static TempState* temp_of(const lsb_ctx* c) {
    for (auto& p : g_temp)
        if (p.first == c) return p.second;
    return nullptr;
}

bool temp_on(const lsb_ctx* c) { return temp_of(c) != nullptr; }

void temp_free(lsb_ctx* c) {
    for (size_t i = 0; i < g_temp.size(); ++i)
        if (g_temp[i].first == c) {
            TempState* t = g_temp[i].second;
            if (t->block) (void)hipFree(t->block);
            if (t->flags) (void)hipFree(t->flags);
            if (t->d_tbar) (void)hipFree(t->d_tbar);
            delete t;
            g_temp.erase(g_temp.begin() + i);
            return;
        }
}

static int temp_alloc(lsb_ctx* c, TempState** out) {
    TempState* t = temp_of(c);
    if (t) {
        *out = t;
        return LSB_OK;
    }
    t = new TempState();
    const size_t bytes = (TS_NSPEC * c->nspec + TP_NPHYS * c->nphys) * sizeof(double);
    int rc = dev_alloc((void**)&t->block, bytes);
    if (rc == LSB_OK) rc = dev_alloc((void**)&t->flags, 16 * sizeof(int));
    if (rc == LSB_OK) rc = dev_alloc((void**)&t->d_tbar, 3 * (size_t)c->nzp * sizeof(double));
    if (rc != LSB_OK) {
        if (t->block) (void)hipFree(t->block);
        if (t->flags) (void)hipFree(t->flags);
        delete t;
        return rc;
    }
    for (int q = 0; q < TS_NSPEC; ++q) t->spec[q] = t->block + (size_t)q * c->nspec;
    for (int q = 0; q < TP_NPHYS; ++q) t->phys[q] = t->block + TS_NSPEC * c->nspec + (size_t)q * c->nphys;
    LSB_HIP(hipMemsetAsync(t->block, 0, bytes, c->stream));
    LSB_HIP(hipMemsetAsync(t->d_tbar, 0, 3 * (size_t)c->nzp * sizeof(double), c->stream));
    t->tbar.assign(c->nzp, 0.0);
    t->tbarp.assign(c->nzp, 0.0);
    t->tbarpp.assign(c->nzp, 0.0);
    g_temp.push_back({c, t});
    *out = t;
    return LSB_OK;
}

int temp_set_params(lsb_ctx* c, const double* tbar, const double* tbarp, const double* tbarpp, double coex,
                    double gx, double gy, double gz) {
    TempState* t;
    LSB_TRY(temp_alloc(c, &t));
    const int n = c->nzp;
    t->tbar.assign(n, 0.0);
    t->tbarp.assign(n, 0.0);
    t->tbarpp.assign(n, 0.0);
    if (tbar) t->tbar.assign(tbar, tbar + n);
    if (tbarp) t->tbarp.assign(tbarp, tbarp + n);
    if (tbarpp) t->tbarpp.assign(tbarpp, tbarpp + n);
    t->coex = coex;
    t->g[0] = gx;
    t->g[1] = gy;
    t->g[2] = gz;
    std::vector<double> h;
    h.insert(h.end(), t->tbar.begin(), t->tbar.end());
    h.insert(h.end(), t->tbarp.begin(), t->tbarp.end());
    h.insert(h.end(), t->tbarpp.begin(), t->tbarpp.end());
    LSB_HIP(hipMemcpyAsync(t->d_tbar, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    LSB_HIP(hipStreamSynchronize(c->stream));
    return LSB_OK;
}

const double* temp_tbar_host(const lsb_ctx* c) {
    TempState* t = temp_of(c);
    return t ? t->tbar.data() : nullptr;
}

const double* temp_tbar_dev(const lsb_ctx* c) {
    TempState* t = temp_of(c);
    return t ? t->d_tbar : nullptr;
}

// which: 0 tempf, 1 tnf, 2 tnmf, 3 nltnf, 4 nltnmf
double* temp_field(lsb_ctx* c, int which) {
    TempState* t = temp_of(c);
    if (!t) return nullptr;
    const int idx[5] = {t->T, t->Tn, t->Tnm, t->Nn, t->Nnm};
    return t->spec[idx[which]];
}

int temp_enable(lsb_ctx* c) {
    TempState* t;
    return temp_alloc(c, &t);
}

// ------------------------------------------------------------------- step
// The temperature part of nlterm_skewsym.  uif: the velocity spectra of the nonlinear term
// (filtered when ftab_fused: the filter still to be applied).
int temp_nlterm(lsb_ctx* c, const double* const* uif, bool filter_u, hipStream_t s) {
    TempState* t = temp_of(c);
    if (!t) return LSB_OK;
    const lsb_config& g = c->cfg;
    const TDims d = tdims(c);
    const long nsp = (long)c->nspec;
    // tempsf = tempf, filtered when dealiasflag >= 2 and some |T| > 1e-10 (navierstokes.f90:1207-1215)
    const double* Tsrc = t->spec[t->T];
    double* Tf = t->spec[TS_TF];
    if (g.dealiasflag >= 2) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, t->flags, 0);
        hipLaunchKernelGGL(k_flag_any_gt, dim3(kFlagBlocks), dim3(256), 0, s, Tsrc, nsp, 1.0e-10, t->flags);
        const double* in[1] = {Tsrc};
        double* out[1] = {t->spec[TS_TFILT]};
        LSB_TRY(sp_specfilter(c, in, out, 1, g.kc, g.filt_order, g.filt_amp, 1, s));
        hipLaunchKernelGGL(k_select, tgrid(nsp), dim3(256), 0, s, t->flags, t->spec[TS_TFILT], Tsrc, Tf, nsp);
        LSB_HIP(hipGetLastError());
    } else {
        LSB_HIP(hipMemcpyAsync(Tf, Tsrc, nsp * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    // velocity of the products (filtered like the momentum products)
    const double* us[3] = {uif[0], uif[1], uif[2]};
    if (filter_u) {
        double* uo[3] = {t->spec[TS_UF], t->spec[TS_UF + 1], t->spec[TS_UF + 2]};
        LSB_TRY(sp_specfilter(c, uif, uo, 3, g.kc, g.filt_order, g.filt_amp, 1, s));
        for (int l = 0; l < 3; ++l) us[l] = uo[l];
    }
    // grad T in spectral space, then T, grad T and u to physical space
    double* dT[3] = {t->spec[TS_D], t->spec[TS_D + 1], t->spec[TS_D + 2]};
    LSB_TRY(sp_dfdx(c, Tf, dT[0], s));
    LSB_TRY(sp_dfdy(c, Tf, dT[1], s));
    {
        const double* in[1] = {Tf};
        double* out[1] = {dT[2]};
        LSB_TRY(zcontract(c, c->d_mdz, c->ldm_z, c->nzp, c->nzp, in, out, 1, s));
    }
    double** P = t->phys;
    {
        const double* in[7] = {Tf, dT[0], dT[1], dT[2], us[0], us[1], us[2]};
        double* out[7] = {P[0], P[1], P[2], P[3], P[8], P[9], P[10]};
        LSB_TRY(horfft_inv(c, c->hp, in, out, 7, c->nzp, s));
    }
    double sum_tbarp = 0.0;
    for (int k = 0; k < c->nzp; ++k) sum_tbarp = sum_tbarp + t->tbarp[k];
    hipLaunchKernelGGL(k_temp_products, tgrid(d.npts), dim3(256), 0, s, d, P[8], P[9], P[10], c->d_paral, P[0], P[1],
                       P[2], P[3], t->d_tbar, sum_tbarp, P[4], P[5], P[6], P[7]);
    LSB_HIP(hipGetLastError());
    // to spectral space: nltempf and F(u_l T); the divergence
    double* N = t->spec[t->NL];
    double* H[3] = {t->spec[TS_H], t->spec[TS_H + 1], t->spec[TS_H + 2]};
    {
        const double* in[4] = {P[4], P[5], P[6], P[7]};
        double* out[4] = {N, H[0], H[1], H[2]};
        LSB_TRY(horfft_fwd(c, c->hp, in, out, 4, c->nzp, s));
    }
    {
        const double* in[1] = {H[2]};
        double* out[1] = {t->spec[TS_DZH]};
        LSB_TRY(zcontract(c, c->d_mdz, c->ldm_z, c->nzp, c->nzp, in, out, 1, s));
    }
    hipLaunchKernelGGL(k_temp_div, tgrid(d.npairs), dim3(256), 0, s, d, c->d_wavx, c->d_wavy,
                       reinterpret_cast<double2*>(N), reinterpret_cast<const double2*>(H[0]),
                       reinterpret_cast<const double2*>(H[1]), reinterpret_cast<const double2*>(t->spec[TS_DZH]));
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// The temperature update (after the step's coefficients are known): new T, rotation, and
// the buoyancy test of the new T.
int temp_update(lsb_ctx* c, hipStream_t s) {
    TempState* t = temp_of(c);
    if (!t) return LSB_OK;
    const lsb_config& g = c->cfg;
    const TDims d = tdims(c);
    const double xkap = g.xnu / g.prtl;
    const double dtkapi = 1. / (xkap * c->dt);
    double sum_abs = 0.0;
    for (double v : t->tbarpp) sum_abs += std::fabs(v);
    double* R = t->spec[t->W];
    hipLaunchKernelGGL(k_temp_rhs, tgrid(2 * d.npairs), dim3(256), 0, s, d, t->spec[t->NL], t->spec[t->Nn],
                       t->spec[t->Nnm], t->spec[t->T], t->spec[t->Tn], t->spec[t->Tnm],
                       sum_abs > 0.0 ? t->d_tbar + 2 * c->nzp : nullptr, c->abfac[0], c->abfac[1], c->abfac[2],
                       c->bdf[0], c->bdf[1], c->bdf[2], c->dt, dtkapi, R);
    LSB_HIP(hipGetLastError());
    {
        double* rh[1] = {R};
        LSB_TRY(helmholtz(c, rh, nullptr, nullptr, 1, dtkapi / c->bdfv, s));
    }
    // nltnmf = nltnf; tnmf = tnf; nltnf = this step's N; tnf = tempf; tempf = rhs
    const int oNnm = t->Nnm, oTnm = t->Tnm;
    t->Nnm = t->Nn;
    t->Nn = t->NL;
    t->NL = oNnm;
    t->Tnm = t->Tn;
    t->Tn = t->T;
    t->T = t->W;
    t->W = oTnm;
    const long nsp = (long)c->nspec;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, t->flags + 1, 0);
    hipLaunchKernelGGL(k_flag_any_gt, dim3(kFlagBlocks), dim3(256), 0, s, t->spec[t->T], nsp, 1.0e-10, t->flags + 1);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int temp_buoyancy(lsb_ctx* c, double* const* rhs, hipStream_t s) {
    TempState* t = temp_of(c);
    if (!t) return LSB_OK;
    if (t->coex == 0.0 && t->g[0] == 0.0 && t->g[1] == 0.0 && t->g[2] == 0.0) return LSB_OK;
    const long nsp = (long)c->nspec;
    hipLaunchKernelGGL(k_buoyancy, tgrid(nsp), dim3(256), 0, s, t->flags + 1, t->spec[t->T], rhs[0], rhs[1], rhs[2],
                       t->coex * t->g[0], t->coex * t->g[1], t->coex * t->g[2], c->cfg.xnu, nsp);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

}  // namespace lsb
