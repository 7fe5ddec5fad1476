// Snapshot and restart records (mpi_io.f90) and the per-step diagnostics of
// the time loop (navierstokes.f90:10-90, 358-511; main.f90:77-84): NaN scan,
// maxcalc extrema / CFL numbers, and the wall skin friction that
// run/compare_results.m:76-79 derives from a snapshot.
//
// All of it is HBM-bound byte work, no arithmetic worth an MFMA:
//   k_snap_pack       one physical x-row (FP64, nxpp stride) -> one FP32 row of
//                     the (x, z, y) record: coalesced on both sides;
//   k_restart_pack    (z, x) transpose of every y plane of a spectral field with
//                     the FP64 -> FP32 rounding, 32 x 32 tiles through LDS;
//   k_restart_unpack  the inverse transpose with the FP32 -> FP64 widening;
//   k_maxcalc, k_nan_scan  grid-stride reductions, one partial per block,
//                     finished on the host in a fixed order (deterministic);
//   k_wall_mean       spanwise mean of two physical planes, 64 x-points per block;
//   k_tavg_accum      the on-the-fly time average, one streaming pass per field.

#include <cfloat>

#include "lsb_internal.h"

namespace lsb {

namespace {

constexpr int kTT = 32;          // transpose tile edge
constexpr int kRedBlocks = 1024;  // partial results of the reductions (4 blocks per CU)

// ----------------------------------------------------------------- snapshot
// mpi_io.f90:411-424: buf(ii, l) = q_l(i, j, k) with ii running x fastest, then
// z, then y.  Block (row, l) copies physical row row = k*ny + j of field l.
// Fields without an additive profile are copied unchanged (no +0.0, which would
// turn -0.0 into +0.0).
__global__ void __launch_bounds__(128) k_snap_pack(FieldList<const double> in, const double* __restrict__ add_u,
                                                   const double* __restrict__ add_t, float* __restrict__ out, int nx,
                                                   int nxpp, int ny, int nzp) {
    const int row = blockIdx.x;
    const int l = blockIdx.y;
    const int k = row / ny, j = row - k * ny;
    const double* src = in.p[l] ? in.p[l] + (size_t)row * nxpp : nullptr;
    float* dst = out + (size_t)l * nx * ny * nzp + ((size_t)j * nzp + k) * nx;
    const double* add = (l == 0) ? add_u : (l == 3) ? add_t : nullptr;
    if (!src) {
        const float a = add ? (float)add[k] : 0.0f;
        for (int i = threadIdx.x; i < nx; i += blockDim.x) dst[i] = a;
    } else if (add) {
        const double a = add[k];
        for (int i = threadIdx.x; i < nx; i += blockDim.x) dst[i] = (float)(src[i] + a);
    } else {
        for (int i = threadIdx.x; i < nx; i += blockDim.x) dst[i] = (float)src[i];
    }
}

// ------------------------------------------------------------------ restart
// mpi_io.f90:47-72: buf(ii) = f(i, j, k) with ii running z fastest, then y, then
// x.  Block (x tile, z tile, field*ny + j); 32 x 8 threads.  A null field writes
// zeros (the temperature slots of a state without temperature).
__global__ void __launch_bounds__(256) k_restart_pack(FieldList<const double> in, float* __restrict__ out, int nx,
                                                      int ny, int nzp) {
    __shared__ float t[kTT][kTT + 1];
    const int i0 = blockIdx.x * kTT, k0 = blockIdx.y * kTT;
    const int l = blockIdx.z / ny, j = blockIdx.z - l * ny;
    const double* f = in.p[l];
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < kTT; r += 8) {
        const int k = k0 + r, i = i0 + tx;
        if (k < nzp && i < nx) t[r][tx] = f ? (float)f[((size_t)k * ny + j) * nx + i] : 0.0f;
    }
    __syncthreads();
    float* o = out + (size_t)l * nx * ny * nzp;
    for (int r = ty; r < kTT; r += 8) {
        const int i = i0 + r, k = k0 + tx;
        if (k < nzp && i < nx) o[((size_t)i * ny + j) * nzp + k] = t[tx][r];
    }
}

// mpi_io.f90:255-285: f(i, j, k) = buf(ii) (the reader's loop); null targets are
// skipped.
__global__ void __launch_bounds__(256) k_restart_unpack(const float* __restrict__ in, FieldList<double> outf, int nx,
                                                        int ny, int nzp) {
    __shared__ float t[kTT][kTT + 1];
    const int i0 = blockIdx.x * kTT, k0 = blockIdx.y * kTT;
    const int l = blockIdx.z / ny, j = blockIdx.z - l * ny;
    double* f = outf.p[l];
    if (!f) return;  // uniform over the block
    const float* src = in + (size_t)l * nx * ny * nzp;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < kTT; r += 8) {
        const int i = i0 + r, k = k0 + tx;
        if (k < nzp && i < nx) t[r][tx] = src[((size_t)i * ny + j) * nzp + k];
    }
    __syncthreads();
    for (int r = ty; r < kTT; r += 8) {
        const int k = k0 + r, i = i0 + tx;
        if (k < nzp && i < nx) f[((size_t)k * ny + j) * nx + i] = (double)t[tx][r];
    }
}

// ------------------------------------------------------------------ maxcalc
struct MaxArgs {
    const double* u;
    const double* v;
    const double* w;
    const double* temp;  // may be null (zero field)
    const double* ubar;  // nzp
    const double* tbar;  // nzp or null
    const double* dz;    // nzp: divisor of plane k's z-CFL factor
    double dt;
    int nx, nxpp, ny, nzp, findex;
};

enum { MX_UMAX, MX_UMIN, MX_VMAX, MX_VMIN, MX_WMAX, MX_WMIN, MX_TMAX, MX_TMIN, MX_UABS, MX_VABS, MX_CFLZ, MX_N };

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// navierstokes.f90:410-450 over the nx physical points of each row.  The z-CFL of
// plane k is (w dt)/dz[k]; plane 0 takes the factor twice (:428-434).
__global__ void __launch_bounds__(256) k_maxcalc(MaxArgs a, double* __restrict__ part) {
    double m[MX_N];
    for (int q = 0; q < MX_TMIN + 1; q += 2) {
        m[q] = -DBL_MAX;    // maxima
        m[q + 1] = DBL_MAX; // minima
    }
    m[MX_UABS] = 0.0;
    m[MX_VABS] = 0.0;
    m[MX_CFLZ] = -DBL_MAX;  // Fortran maxval of an empty section (findex = 0) is -huge
    const int nrows = a.ny * a.nzp;
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const int k = row / a.ny;
        const size_t base = (size_t)row * a.nxpp;
        const double ub = a.ubar[k], dzk = a.dz[k];
        const double tb = a.tbar ? a.tbar[k] : 0.0;
        for (int i = threadIdx.x; i < a.nx; i += blockDim.x) {
            const double u = a.u[base + i] + ub;
            const double v = a.v[base + i];
            const double w = a.w[base + i];
            const double tp = (a.temp ? a.temp[base + i] : 0.0) + tb;
            m[MX_UMAX] = dmax(m[MX_UMAX], u);
            m[MX_UMIN] = dmin(m[MX_UMIN], u);
            m[MX_VMAX] = dmax(m[MX_VMAX], v);
            m[MX_VMIN] = dmin(m[MX_VMIN], v);
            m[MX_WMAX] = dmax(m[MX_WMAX], w);
            m[MX_WMIN] = dmin(m[MX_WMIN], w);
            m[MX_TMAX] = dmax(m[MX_TMAX], tp);
            m[MX_TMIN] = dmin(m[MX_TMIN], tp);
            m[MX_UABS] = dmax(m[MX_UABS], fabs(u));
            m[MX_VABS] = dmax(m[MX_VABS], fabs(v));
            if (i < a.findex) {
                double ws = (w * a.dt) / dzk;
                if (k == 0) ws = (ws * a.dt) / dzk;
                m[MX_CFLZ] = dmax(m[MX_CFLZ], fabs(ws));
            }
        }
    }
    // wave then block reduction
    for (int q = 0; q < MX_N; ++q) {
        const bool mn = (q == MX_UMIN || q == MX_VMIN || q == MX_WMIN || q == MX_TMIN);
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_xor(m[q], o);
            m[q] = mn ? dmin(m[q], y) : dmax(m[q], y);
        }
    }
    __shared__ double sh[4][MX_N];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int q = 0; q < MX_N; ++q) sh[wave][q] = m[q];
    __syncthreads();
    if (threadIdx.x < MX_N) {
        const int q = threadIdx.x;
        const bool mn = (q == MX_UMIN || q == MX_VMIN || q == MX_WMIN || q == MX_TMIN);
        double r = sh[0][q];
        for (int wv = 1; wv < (int)(blockDim.x >> 6); ++wv) r = mn ? dmin(r, sh[wv][q]) : dmax(r, sh[wv][q]);
        part[(size_t)blockIdx.x * MX_N + q] = r;
    }
}

// any(isnan(f)) of nf fields of n reals: part[block] = 1 when the block saw a NaN.
__global__ void __launch_bounds__(256) k_nan_scan(FieldList<const double> f, int nf, size_t n,
                                                  double* __restrict__ part) {
    int bad = 0;
    for (int l = 0; l < nf; ++l) {
        const double* p = f.p[l];
        for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
            bad |= (p[t] != p[t]);
    }
    const int any = __any(bad);
    __shared__ int sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = any;
    __syncthreads();
    if (threadIdx.x == 0) {
        int r = 0;
        for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) r |= sh[wv];
        part[blockIdx.x] = r ? 1.0 : 0.0;
    }
}

// spanwiseavg.m:3 on two physical planes (2, ny, nxpp): out[p*nx + i] = sum_j / ny.
// Block (64 x-points, plane p), four y-groups of one wave each summing j = g, g+4, ...;
// the four partial sums are added in a fixed order.
__global__ void __launch_bounds__(256) k_wall_mean(const double* __restrict__ ph, int nx, int nxpp, int ny,
                                                   double* __restrict__ out) {
    __shared__ double sh[4][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane, p = blockIdx.y;
    double s = 0.0;
    if (i < nx)
        for (int j = g; j < ny; j += 4) s += ph[((size_t)p * ny + j) * nxpp + i];
    sh[g][lane] = s;
    __syncthreads();
    if (g == 0 && i < nx) out[p * nx + i] = (((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane]) / (double)ny;
}

// orrsomm.f90:650-656: per plane k, sum over the nx*ny physical points of
// 0.5 (u^2 + v^2 + w^2), u + ubar(k) when ubar is given.  Block (k, part).
constexpr int kKeParts = 8;
__global__ void __launch_bounds__(256) k_plane_ke(const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, const double* __restrict__ ubar,
                                                  int nx, int nxpp, int ny, double* __restrict__ part) {
    const int k = blockIdx.x, p = blockIdx.y;
    const long per_plane = (long)ny * nx;
    const double ub = ubar ? ubar[k] : 0.0;
    double s = 0.0;
    for (long e = (long)p * blockDim.x + threadIdx.x; e < per_plane; e += (long)kKeParts * blockDim.x) {
        const int j = (int)(e / nx), i = (int)(e - (long)j * nx);
        const long off = ((long)k * ny + j) * nxpp + i;
        const double a = u[off] + ub, b = v[off], cc = w[off];
        s += 0.5 * (a * a + b * b + cc * cc);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(long)k * kKeParts + p] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// the (kx, ky) = (0, 0) mode (two reals at the start of every plane) set to zero
__global__ void k_zero_mean(FieldList<double> f, int nf, int nzp, long plane) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * nzp * 2) return;
    const int l = t / (nzp * 2), r = t - l * nzp * 2;
    f.p[l][(long)(r >> 1) * plane + (r & 1)] = 0.0;
}

// main.f90:69-76: avg_l = avg_l + f_l / div over n2 complex pairs of field l =
// blockIdx.y.  A division, as the reference's q/iomod (not a product with 1/div,
// which rounds differently).
__global__ void __launch_bounds__(256) k_tavg_accum(FieldList<const double> f, FieldList<double> avg, size_t n2,
                                                    double div) {
    const double2* __restrict__ src = reinterpret_cast<const double2*>(f.p[blockIdx.y]);
    double2* __restrict__ dst = reinterpret_cast<double2*>(avg.p[blockIdx.y]);
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x) {
        const double2 q = src[t];
        double2 a = dst[t];
        a.x = a.x + q.x / div;
        a.y = a.y + q.y / div;
        dst[t] = a;
    }
}

// derivatives.f90:9-70 zinteg: cumulative z integral of each of np pencils (Fortran
// (n1, n2, nzp): a pencil's values np apart, so consecutive threads read consecutive
// memory).  w != null: per = 1, tint(k) = sum over k' = k..nz of w(k') tmp(k') with the
// per-point weights w; otherwise the trapezoid rule on zpts, from the top (rev = 0) or
// from the wall (rev = 1).  0-based k here.
__global__ void __launch_bounds__(256) k_zinteg(const double* __restrict__ q, double* __restrict__ out, long np,
                                                int nzp, const double* __restrict__ zp,
                                                const double* __restrict__ w, int mapping, int rev) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    double integ = 0.0;
    if (w) {
        out[(long)(nzp - 1) * np + t] = 0.0;
        for (int k = nzp - 2; k >= 1; --k) {
            integ = integ + w[k] * q[(long)k * np + t];
            out[(long)k * np + t] = integ;
        }
        out[t] = mapping == 1 ? out[np + t] : 0.0;
    } else if (rev == 0) {
        out[(long)(nzp - 1) * np + t] = 0.0;
        for (int k = nzp - 1; k >= 1; --k) {
            integ = integ + 0.5 * (q[(long)(k - 1) * np + t] + q[(long)k * np + t]) * (zp[k - 1] - zp[k]);
            out[(long)(k - 1) * np + t] = integ;
        }
        if (mapping == 1) out[t] = out[np + t];
    } else {
        out[t] = 0.0;
        for (int k = 1; k < nzp; ++k) {
            integ = integ + 0.5 * (q[(long)(k - 1) * np + t] + q[(long)k * np + t]) * (zp[k - 1] - zp[k]);
            out[(long)k * np + t] = integ;
        }
    }
}

int diag_scratch(lsb_ctx* c, double** p) {
    if (!c->d_diag_red) LSB_TRY(dev_alloc((void**)&c->d_diag_red, (size_t)kDiagDoubles * sizeof(double)));
    *p = c->d_diag_red;
    return LSB_OK;
}

}  // namespace

int snap_pack(lsb_ctx* c, const double* const* phys, const double* add_u, const double* add_t, float* out,
              hipStream_t s) {
    FieldList<const double> in{};
    for (int l = 0; l < 5; ++l) in.p[l] = phys[l];
    const dim3 grid((unsigned)(c->ny * c->nzp), 5u);
    hipLaunchKernelGGL(k_snap_pack, grid, dim3(128), 0, s, in, add_u, add_t, out, c->nx, c->nxpp, c->ny, c->nzp);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int restart_pack(lsb_ctx* c, const double* const* spec, int nf, float* out, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields || (long)nf * c->ny > 65535) return LSB_E_ARG;
    FieldList<const double> in{};
    for (int l = 0; l < nf; ++l) in.p[l] = spec[l];
    const dim3 grid((unsigned)((c->nx + kTT - 1) / kTT), (unsigned)((c->nzp + kTT - 1) / kTT), (unsigned)(nf * c->ny));
    hipLaunchKernelGGL(k_restart_pack, grid, dim3(kTT, 8), 0, s, in, out, c->nx, c->ny, c->nzp);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int restart_unpack(lsb_ctx* c, const float* in, double* const* spec, int nf, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields || (long)nf * c->ny > 65535) return LSB_E_ARG;
    FieldList<double> o{};
    for (int l = 0; l < nf; ++l) o.p[l] = spec[l];
    const dim3 grid((unsigned)((c->nx + kTT - 1) / kTT), (unsigned)((c->nzp + kTT - 1) / kTT), (unsigned)(nf * c->ny));
    hipLaunchKernelGGL(k_restart_unpack, grid, dim3(kTT, 8), 0, s, in, o, c->nx, c->ny, c->nzp);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int maxcalc_reduce(lsb_ctx* c, const double* u, const double* v, const double* w, const double* temp,
                   const double* tbar_host, double dt, int findex, double* stats, hipStream_t s) {
    const int nzp = c->nzp, nz = c->nz;
    if (kRedBlocks * MX_N + 2 * nzp > kDiagDoubles) {
        set_error("maxcalc: nzp too large for the reduction scratch");
        return LSB_E_ARG;
    }
    double* scr;
    LSB_TRY(diag_scratch(c, &scr));
    // navierstokes.f90:428-437 divisors: zpts(k) - zpts(k+1) for k = 1..nz (plane 1 twice),
    // zpts(nz) - zpts(nzp) for plane nzp
    std::vector<double> dz(nzp);
    for (int k = 0; k < nz; ++k) dz[k] = c->zpts[k] - c->zpts[k + 1];
    dz[nz] = c->zpts[nz - 1] - c->zpts[nz];
    double* d_dz = scr + kRedBlocks * MX_N;
    double* d_tbar = d_dz + nzp;
    LSB_HIP(hipMemcpyAsync(d_dz, dz.data(), nzp * sizeof(double), hipMemcpyHostToDevice, s));
    if (tbar_host) LSB_HIP(hipMemcpyAsync(d_tbar, tbar_host, nzp * sizeof(double), hipMemcpyHostToDevice, s));
    MaxArgs a{u, v, w, temp, c->d_paral, tbar_host ? d_tbar : nullptr, d_dz, dt, c->nx, c->nxpp, c->ny, nzp, findex};
    hipLaunchKernelGGL(k_maxcalc, dim3(kRedBlocks), dim3(256), 0, s, a, scr);
    LSB_HIP(hipGetLastError());
    std::vector<double> part((size_t)kRedBlocks * MX_N);
    LSB_HIP(hipMemcpyAsync(part.data(), scr, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    LSB_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < MX_N; ++q) {
        const bool mn = (q == MX_UMIN || q == MX_VMIN || q == MX_WMIN || q == MX_TMIN);
        double r = part[q];
        for (int b = 1; b < kRedBlocks; ++b) {
            const double y = part[(size_t)b * MX_N + q];
            r = mn ? std::min(r, y) : std::max(r, y);
        }
        stats[q] = r;
    }
    return LSB_OK;
}

int nan_scan(lsb_ctx* c, const double* const* f, int nf, size_t n, bool* found, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields) return LSB_E_ARG;
    double* scr;
    LSB_TRY(diag_scratch(c, &scr));
    FieldList<const double> in{};
    for (int l = 0; l < nf; ++l) in.p[l] = f[l];
    hipLaunchKernelGGL(k_nan_scan, dim3(kRedBlocks), dim3(256), 0, s, in, nf, n, scr);
    LSB_HIP(hipGetLastError());
    std::vector<double> part(kRedBlocks);
    LSB_HIP(hipMemcpyAsync(part.data(), scr, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, s));
    LSB_HIP(hipStreamSynchronize(s));
    *found = false;
    for (double v : part) *found = *found || (v != 0.0);
    return LSB_OK;
}

int plane_ke(lsb_ctx* c, const double* u, const double* v, const double* w, bool add_ubar, double* ee_host,
             hipStream_t s) {
    if ((long)c->nzp * kKeParts > kDiagDoubles) {
        set_error("ener: nzp too large for the reduction scratch");
        return LSB_E_ARG;
    }
    double* scr;
    LSB_TRY(diag_scratch(c, &scr));
    hipLaunchKernelGGL(k_plane_ke, dim3((unsigned)c->nzp, (unsigned)kKeParts), dim3(256), 0, s, u, v, w,
                       add_ubar ? c->d_paral : nullptr, c->nx, c->nxpp, c->ny, scr);
    LSB_HIP(hipGetLastError());
    std::vector<double> part((size_t)c->nzp * kKeParts);
    LSB_HIP(hipMemcpyAsync(part.data(), scr, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    LSB_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < c->nzp; ++k) {
        double e = 0.0;
        for (int p = 0; p < kKeParts; ++p) e += part[(size_t)k * kKeParts + p];
        ee_host[k] = e;
    }
    return LSB_OK;
}

int zero_mean_modes(lsb_ctx* c, double* const* spec, int nf, hipStream_t s) {
    FieldList<double> f{};
    for (int l = 0; l < nf; ++l) f.p[l] = spec[l];
    const int n = nf * c->nzp * 2;
    hipLaunchKernelGGL(k_zero_mean, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, nf, c->nzp,
                       (long)c->plane_s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int tavg_accum(lsb_ctx* c, const double* const* f, double* const* avg, int nf, int div, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields || div < 1 || (c->nspec & 1)) return LSB_E_ARG;
    FieldList<const double> in{};
    FieldList<double> out{};
    for (int l = 0; l < nf; ++l) {
        in.p[l] = f[l];
        out.p[l] = avg[l];
    }
    hipLaunchKernelGGL(k_tavg_accum, dim3(1024u, (unsigned)nf), dim3(256), 0, s, in, out, c->nspec / 2, (double)div);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int zinteg_dev(lsb_ctx* c, const double* tmp, double* tint, int n1, int n2, int mapping, int per, int rev,
               const double* w, hipStream_t s) {
    const long np = (long)n1 * n2;
    hipLaunchKernelGGL(k_zinteg, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, tmp, tint, np, c->nzp, c->d_zpts,
                       per == 1 ? w : nullptr, mapping, rev);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int wall_mean(lsb_ctx* c, const double* phys2, double* mean_host, hipStream_t s) {
    if (2 * c->nx > kDiagDoubles) {
        set_error("wall_shear: nx too large for the reduction scratch");
        return LSB_E_ARG;
    }
    double* scr;
    LSB_TRY(diag_scratch(c, &scr));
    hipLaunchKernelGGL(k_wall_mean, dim3((unsigned)((c->nx + 63) / 64), 2u), dim3(256), 0, s, phys2, c->nx, c->nxpp,
                       c->ny, scr);
    LSB_HIP(hipGetLastError());
    LSB_HIP(hipMemcpyAsync(mean_host, scr, 2 * c->nx * sizeof(double), hipMemcpyDeviceToHost, s));
    LSB_HIP(hipStreamSynchronize(s));
    return LSB_OK;
}

}  // namespace lsb
