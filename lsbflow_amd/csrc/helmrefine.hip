// Batched Helmholtz solves in four precisions (helmholtz.f90:130-208, homogeneous
// Dirichlet rows): every (kx, ky) pencil of nf spectral fields, interior rows 2..nz,
//   (D2 - k^2 - sigma) u = rhs,   u = E diag(1/(lambda - k^2 - sigma)) E^-1 rhs,
// u(1) = u(nzp) = 0, written over rhs.
//   LSB_HELM_FP64   the FP64 projections (the step's solver);
//   LSB_HELM_FP32   both projections on the FP32 matrix cores (E, E^-1 rounded to FP32,
//                   the mode values in FP32), the result widened to FP64;
//   LSB_HELM_REFINE the FP32 solve, then one FP64 refinement: the residual
//                   r = rhs - (D2 u - (k^2 + sigma) u) in FP64 with the D2 = d.d the
//                   eigen-decomposition was made of, a second FP32 solve of r, u += du;
//   LSB_HELM_SPLIT32 the mixed form's velocity solve (zops.hip k_zhelm32): the FP64
//                   leading term -rhs/(k^2 + sigma) and the FP32 projections of the
//                   correction, used when sigma >= 16 max|lambda| (else the FP64 solve).
// The mode values, the FP64 operator and the residual live in the context's staging
// slots 20-22 (no other entry point uses them).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "lsb_internal.h"

namespace lsb {

// t[k][col] /= (lambda_k - kx^2 - ky^2 - sigma) for the nzm mode rows of one field
__global__ void k_helm_modescale(float* __restrict__ t, const double* __restrict__ lam,
                                 const double* __restrict__ wavx, const double* __restrict__ wavy, int nx, int ny,
                                 int nzm, double sigma) {
    const long P = (long)nx * ny;
    const long n = P * nzm;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e / P);
        const long col = e - (long)k * P;
        const int i = (int)(col % nx), j = (int)(col / nx);
        const double den = lam[k] - wavx[i] * wavx[i] - wavy[j] * wavy[j] - sigma;
        t[e] = (float)((double)t[e] / den);
    }
}

// r = aux - D2 u on the interior rows (from the contraction) plus (k^2 + sigma) u
__global__ void k_helm_resid(double* __restrict__ r, const double* __restrict__ u, const double* __restrict__ wavx,
                             const double* __restrict__ wavy, int nx, int ny, int nzm, double sigma) {
    const long P = (long)nx * ny;
    const long n = P * nzm;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long col = e % P;
        const int i = (int)(col % nx), j = (int)(col / nx);
        const double k2 = wavx[i] * wavx[i] + wavy[j] * wavy[j];
        r[e] = r[e] + (k2 + sigma) * u[P + e];  // u interior row k + 1
    }
}

// u[interior] += du[interior]; boundary rows zero
__global__ void k_helm_update(double* __restrict__ u, const double* __restrict__ du, long P, int nzp) {
    const long n = P * nzp;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e / P);
        u[e] = (k == 0 || k == nzp - 1) ? 0.0 : u[e] + du[e];
    }
}

__global__ void k_zero_rows(double* __restrict__ u, long P, int nzp) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < P; e += (long)gridDim.x * blockDim.x) {
        u[e] = 0.0;
        u[(long)(nzp - 1) * P + e] = 0.0;
    }
}

static dim3 grid_for(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 16) b = 65535L * 16;
    return dim3((unsigned)b);
}

// FP32 solve of one field: interior rows of `in` (FP64, nzp rows) -> interior rows of
// `out` (FP64); `t` holds nzm FP32 mode rows.  in and out may be the same array.
static int fp32_solve(lsb_ctx* c, const double* in, double* out, float* t, double sigma, hipStream_t s) {
    const int nzm = c->nzm, nzp = c->nzp;
    const long P = (long)c->plane_s;
    const int ldh = ((nzm + 63) / 64) * 64 + 64;
    const double* rin[1] = {in + P};
    float* tm[1] = {t};
    LSB_TRY((zcontract_f32<double, float>(c, c->d_simat, ldh, nzm, nzm, rin, tm, 1, s, P, nullptr)));
    hipLaunchKernelGGL(k_helm_modescale, grid_for(P * nzm), dim3(256), 0, s, t, c->d_diag, c->d_wavx, c->d_wavy,
                       c->nx, c->ny, nzm, sigma);
    LSB_HIP(hipGetLastError());
    const float* tin[1] = {t};
    double* uo[1] = {out + P};
    LSB_TRY((zcontract_f32<float, double>(c, c->d_smat, ldh, nzm, nzm, tin, uo, 1, s, P, nullptr)));
    hipLaunchKernelGGL(k_zero_rows, grid_for(P), dim3(256), 0, s, out, P, nzp);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int helmholtz_precision(lsb_ctx* c, double* rhs, int nf, double sigma, int mode, hipStream_t s) {
    if (mode == LSB_HELM_FP64 || mode == LSB_HELM_SPLIT32) {
        for (int f0 = 0; f0 < nf; f0 += kMaxFields) {
            const int nb = std::min(kMaxFields, nf - f0);
            double* r[kMaxFields];
            for (int q = 0; q < nb; ++q) r[q] = rhs + (size_t)(f0 + q) * c->nspec;
            LSB_TRY(helmholtz(c, r, nullptr, nullptr, nb, sigma, s, mode == LSB_HELM_SPLIT32));
        }
        return LSB_OK;
    }
    if (mode != LSB_HELM_FP32 && mode != LSB_HELM_REFINE) {
        set_error("solve_helmholtz_mixed: mode must be LSB_HELM_FP64, LSB_HELM_FP32, LSB_HELM_REFINE or "
                  "LSB_HELM_SPLIT32");
        return LSB_E_ARG;
    }
    const int nzm = c->nzm, nzp = c->nzp;
    const long P = (long)c->plane_s;
    void* tv = nullptr;
    LSB_TRY(stage_get(c, 20, (size_t)P * nzm * sizeof(float), &tv));
    float* t = (float*)tv;
    double *res = nullptr, *d2 = nullptr;
    int ld2 = 0;
    if (mode == LSB_HELM_REFINE) {
        // FP64 operator rows 2..nz of D2 = d.d (the matrix the eigen-decomposition factored),
        // padded for the contraction; the residual / correction field
        ld2 = ((nzm + 63) / 64) * 64 + 64;
        std::vector<double> h((size_t)ld2 * ((nzp + 3) / 4 * 4), 0.0);
        for (int j = 0; j < nzp; ++j)
            for (int i = 0; i < nzm; ++i) h[(size_t)j * ld2 + i] = c->d2[(size_t)j * nzp + (i + 1)];
        void* dv = nullptr;
        LSB_TRY(stage_get(c, 21, h.size() * sizeof(double), &dv));
        d2 = (double*)dv;
        LSB_HIP(hipMemcpyAsync(d2, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
        // the copy reads h before the function returns
        LSB_HIP(hipStreamSynchronize(s));
        void* rv = nullptr;
        LSB_TRY(stage_get(c, 22, (size_t)P * nzp * sizeof(double), &rv));
        res = (double*)rv;
    }
    for (int f = 0; f < nf; ++f) {
        double* u = rhs + (size_t)f * c->nspec;
        if (mode == LSB_HELM_FP32) {
            LSB_TRY(fp32_solve(c, u, u, t, sigma, s));
            continue;
        }
        // res (interior rows) = rhs; the first solve in place
        LSB_HIP(hipMemcpyAsync(res + P, u + P, (size_t)P * nzm * sizeof(double), hipMemcpyDeviceToDevice, s));
        LSB_TRY(fp32_solve(c, u, u, t, sigma, s));
        // res = rhs - D2 u + (k^2 + sigma) u on the interior rows
        const double* uin[1] = {u};
        double* ro[1] = {res + P};
        const double* ax[1] = {res + P};
        LSB_TRY(zcontract_ex(c, d2, ld2, nzm, nzp, uin, ro, 1, P, ax, s));
        hipLaunchKernelGGL(k_helm_resid, grid_for(P * nzm), dim3(256), 0, s, res + P, u, c->d_wavx, c->d_wavy,
                           c->nx, c->ny, nzm, sigma);
        LSB_HIP(hipGetLastError());
        // du = FP32 solve of the residual (in place in res), u += du
        LSB_TRY(fp32_solve(c, res, res, t, sigma, s));
        hipLaunchKernelGGL(k_helm_update, grid_for(P * nzp), dim3(256), 0, s, u, res, P, nzp);
        LSB_HIP(hipGetLastError());
    }
    return LSB_OK;
}

}  // namespace lsb

using namespace lsb;

extern "C" int lsb_solve_helmholtz_mixed_dev(lsb_ctx* c, double* rhs, double dtnui, int nfields, int mode,
                                             void* stream) {
    if (!c || !rhs || nfields < 1) {
        set_error("solve_helmholtz_mixed_dev: null context or array, or nfields < 1");
        return LSB_E_ARG;
    }
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return helmholtz_precision(c, rhs, nfields, dtnui, mode, s);
}
