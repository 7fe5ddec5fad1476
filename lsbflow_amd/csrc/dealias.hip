// 3/2-rule (or 5/3, 8/5, 2x) padded product, dealiasing.f90:87-447 at one rank.
//
//   a, b physical -> horfft(-1) -> embed into the padded spectrum ->
//   padded inverse transform -> product -> padded forward transform (scale
//   1/(nx_lg ny_lg)) -> truncate -> horfft(+1) -> ab physical.
//
// Placement of the ky modes in the padded spectrum:
//   ref_compat = 1: ky slot j -> j for all j (dealiasing.f90:329-330, 435), i.e.
//                   the reference's layout, which treats the negative-ky half as
//                   positive frequencies of the padded grid;
//   ref_compat = 0: negative ky (j >= ny/2) -> j - ny + ny_lg (textbook 3/2 rule).
// The kx modes 0..nx/2-1 go to the first slots of each padded row in both modes.

#include "lsb_internal.h"

namespace lsb {

__global__ void k_embed(const double2* __restrict__ in, double2* __restrict__ out, int lh, int ny, int lh_lg,
                        int ny_lg, int nzp, int ref) {
    const long n = (long)nzp * ny * lh;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int m = (int)(t % lh);
        const long jk = t / lh;
        const int j = (int)(jk % ny);
        const int k = (int)(jk / ny);
        const int jd = (ref || j < ny / 2) ? j : j - ny + ny_lg;
        out[((long)k * ny_lg + jd) * lh_lg + m] = in[t];
    }
}

__global__ void k_truncate(const double2* __restrict__ in, double2* __restrict__ out, int lh, int ny, int lh_lg,
                           int ny_lg, int nzp, int ref) {
    const long n = (long)nzp * ny * lh;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int m = (int)(t % lh);
        const long jk = t / lh;
        const int j = (int)(jk % ny);
        const int k = (int)(jk / ny);
        const int jd = (ref || j < ny / 2) ? j : j - ny + ny_lg;
        out[t] = in[((long)k * ny_lg + jd) * lh_lg + m];
    }
}

__global__ void k_mul(double* __restrict__ a, const double* __restrict__ b, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        a[t] = a[t] * b[t];
}

static dim3 grid_for(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 8) b = 65535L * 8;
    return dim3((unsigned)b);
}

int dealias_setup(lsb_ctx* c) {
    if (c->nx_lg) return LSB_OK;
    const int nx = c->nx, ny = c->ny;
    int nxl, nyl;
    // dealiasing.f90:134-154 with nproch = 1
    if ((nx % 2) == 0 && nx % 4 == 0)
        nxl = 3 * nx / 2;
    else if (nx % 3 == 0)
        nxl = 5 * nx / 3;
    else if (nx % 5 == 0)
        nxl = 8 * nx / 5;
    else
        nxl = 2 * nx;
    if (ny % 2 == 0)
        nyl = 3 * ny / 2;
    else if (ny % 3 == 0)
        nyl = 5 * ny / 3;
    else if (ny % 5 == 0)
        nyl = 8 * ny / 5;
    else
        nyl = 2 * ny;
    LSB_TRY(hor_plan_create(&c->hp_lg, nxl, nyl));
    c->nx_lg = nxl;
    c->ny_lg = nyl;
    return LSB_OK;
}

// dealiasing.f90:449-618 dealiasf_2d_fftw: the padded product of two spectral fields,
// spectral in and out -- embed in the 3/2 (or larger) spectrum, inverse transforms,
// product, forward transform (scaled 1/(nx_lg ny_lg)), truncation.
int dealias_product_spec(lsb_ctx* c, const double* af, const double* bf, double* abf, hipStream_t s) {
    LSB_TRY(dealias_setup(c));
    const int nzp = c->nzp, lh = c->nx / 2, ny = c->ny;
    const int nxl = c->nx_lg, nyl = c->ny_lg, lhl = nxl / 2;
    const size_t nspec_lg = (size_t)nxl * nyl * nzp;
    const size_t nphys_lg = (size_t)(nxl + 2) * nyl * nzp;
    double *al, *bl, *pa, *pb;
    LSB_TRY(stage_get(c, 12, nspec_lg * sizeof(double), (void**)&al));
    LSB_TRY(stage_get(c, 13, nspec_lg * sizeof(double), (void**)&bl));
    LSB_TRY(stage_get(c, 14, nphys_lg * sizeof(double), (void**)&pa));
    LSB_TRY(stage_get(c, 15, nphys_lg * sizeof(double), (void**)&pb));
    const int ref = c->cfg.dealias_ref_compat ? 1 : 0;
    LSB_HIP(hipMemsetAsync(al, 0, nspec_lg * sizeof(double), s));
    LSB_HIP(hipMemsetAsync(bl, 0, nspec_lg * sizeof(double), s));
    const long ne = (long)nzp * ny * lh;
    hipLaunchKernelGGL(k_embed, grid_for(ne), dim3(256), 0, s, (const double2*)af, (double2*)al, lh, ny, lhl, nyl, nzp,
                       ref);
    hipLaunchKernelGGL(k_embed, grid_for(ne), dim3(256), 0, s, (const double2*)bf, (double2*)bl, lh, ny, lhl, nyl, nzp,
                       ref);
    LSB_HIP(hipGetLastError());
    {
        const double* in[2] = {al, bl};
        double* out[2] = {pa, pb};
        LSB_TRY(horfft_inv(c, c->hp_lg, in, out, 2, nzp, s));
    }
    hipLaunchKernelGGL(k_mul, grid_for((long)nphys_lg), dim3(256), 0, s, pa, pb, (long)nphys_lg);
    LSB_HIP(hipGetLastError());
    {
        const double* in[1] = {pa};
        double* out[1] = {al};
        LSB_TRY(horfft_fwd(c, c->hp_lg, in, out, 1, nzp, s));
    }
    hipLaunchKernelGGL(k_truncate, grid_for(ne), dim3(256), 0, s, (const double2*)al, (double2*)abf, lh, ny, lhl, nyl,
                       nzp, ref);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// dealiasing.f90:262-447 dealiasing_2d_fftw: physical in and out, the spectral product
// between a forward and an inverse 2-D transform.
int dealias_product(lsb_ctx* c, const double* a, const double* b, double* ab, hipStream_t s) {
    LSB_TRY(dealias_setup(c));
    double *af, *bf;
    LSB_TRY(stage_get(c, 10, c->nspec * sizeof(double), (void**)&af));
    LSB_TRY(stage_get(c, 11, c->nspec * sizeof(double), (void**)&bf));
    {
        const double* in[2] = {a, b};
        double* out[2] = {af, bf};
        LSB_TRY(horfft_fwd(c, c->hp, in, out, 2, c->nzp, s));
    }
    LSB_TRY(dealias_product_spec(c, af, bf, af, s));
    {
        const double* in[1] = {af};
        double* out[1] = {ab};
        LSB_TRY(horfft_inv(c, c->hp, in, out, 1, c->nzp, s));
    }
    return LSB_OK;
}

}  // namespace lsb
