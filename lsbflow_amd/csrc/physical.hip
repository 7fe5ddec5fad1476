// Physical-space kernels: skew-symmetric products, SGS models, 3-point filters.
//
// Physical layout: [nzp][ny][nxpp] reals; only i < nx is computed (pads are
// written as zero).  Tensors are addressed through a list of 9 field pointers,
// element (l, m) at index l + 3*m like the reference's (..., 3, 3) arrays; the
// step passes aliased pointers for symmetric tensors (same values, fewer
// fields).

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "lsb_internal.h"
#include "physical.h"

namespace lsb {

struct PhDims {
    long npts;  // nzp*ny*nx
    int nx, ny, nxpp, nzp;
};

static PhDims ph_dims(const lsb_ctx* c) {
    PhDims d;
    d.nx = c->nx;
    d.ny = c->ny;
    d.nxpp = c->nxpp;
    d.nzp = c->nzp;
    d.npts = (long)c->nzp * c->ny * c->nx;
    return d;
}

static dim3 ph_grid(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 8) b = 65535L * 8;
    return dim3((unsigned)b);
}

#define PH_LOOP(d)                                                                                          \
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (d).npts; e += (long)gridDim.x * blockDim.x)

__device__ __forceinline__ void ph_index(const PhDims& d, long e, int& i, int& j, int& k, long& off) {
    i = (int)(e % d.nx);
    const long jk = e / d.nx;
    j = (int)(jk % d.ny);
    k = (int)(jk / d.ny);
    off = ((long)k * d.ny + j) * d.nxpp + i;
}

// Zero the two pad reals of every x-row of nf fields.
__global__ void k_zero_pads(FieldList<double> f, int nf, long nrows, int nx, int nxpp) {
    const long n = nrows * nf * 2;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int q = (int)(t & 1);
        const long r = (t >> 1) % nrows;
        const int fi = (int)((t >> 1) / nrows);
        f.p[fi][r * nxpp + nx + q] = 0.0;
    }
}

// ---------------------------------------------------------- SGS per point
// sgsterm.f90:283-423 sigma model: eddy viscosity from the singular values of g,
// through the invariants of G = g^T g and the trigonometric roots of its
// characteristic cubic.  The same quantities as the reference with fewer
// operations: G is symmetric (6 dot products), tr(G^2) is the sum of squares of
// G, the three roots use cos(pi/3 +- a) = cos(a)/2 -+ (sqrt 3/2) sin(a) from one
// sincos (a in [0, pi/3]), and the constant divisions are multiplications.
__device__ __forceinline__ double sigma_nu(const double g[3][3], double delta) {
    const double pi = 3.141592653589793;
    double G[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) G[a][b] = g[0][a] * g[0][b] + g[1][a] * g[1][b] + g[2][a] * g[2][b];
    G[1][0] = G[0][1];
    G[2][0] = G[0][2];
    G[2][1] = G[1][2];
    const double i1 = G[0][0] + G[1][1] + G[2][2];
    const double trg2 = G[0][0] * G[0][0] + G[1][1] * G[1][1] + G[2][2] * G[2][2] +
                        2.0 * (G[0][1] * G[0][1] + G[0][2] * G[0][2] + G[1][2] * G[1][2]);
    const double i2 = 0.5 * (i1 * i1 - trg2);
    const double i3 = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                      G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                      G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    const double third = 1.0 / 3.0;
    double a1 = i1 * i1 * (1.0 / 9.0) - i2 * third;
    const double a2 = i1 * i1 * i1 * (1.0 / 27.0) - i1 * i2 * (1.0 / 6.0) + 0.5 * i3;
    if (a1 < 0.0) a1 = fabs(a1);
    const double sa = sqrt(a1);
    const double pre = a2 / (sa * sa * sa);
    double a3;
    if (pre >= 1. || isnan(pre))
        a3 = 0.0;
    else if (pre <= -1.)
        a3 = pi / 3.;
    else
        a3 = acos(pre) * third;
    double sn, cs;
    sincos(a3, &sn, &cs);
    const double r3 = 0.86602540378443865;  // sqrt(3)/2
    const double c1 = cs;                   // cos(a3)
    const double c2 = 0.5 * cs - r3 * sn;   // cos(pi/3 + a3)
    const double c3 = 0.5 * cs + r3 * sn;   // cos(pi/3 - a3)
    const double t = i1 * third, s2a = 2. * sa;
    const double s1 = sqrt(fabs(t + s2a * c1));
    const double s2 = sqrt(fabs(t - s2a * c2));
    const double s3 = sqrt(fabs(t - s2a * c3));
    double ds = s3 * (s1 - s2) * (s2 - s3);
    if (s1 != 0.) ds = ds / (s1 * s1);
    if (ds < 0.0 || isnan(ds)) ds = 0.0;
    const double cd = 1.35 * delta;
    return -(cd * cd) * ds;
}

// ---------------------------------------------------------- nonlinear term
// navierstokes.f90:725-769: base flow added, nlt_l = -1/2 sum_m u_m du_l/dx_m,
// uu_lm = u_l u_m, tau from the SGS model, h_lm = 1/2 uu_lm + tau_lm.
// Arithmetic in FP64 whatever the storage type T.
// Element access by byte offset: with a 32-bit offset the loads and stores take the
// scalar-base + vector-offset form (no 64-bit address arithmetic per field).
template <class T, class Off>
__device__ __forceinline__ T ldb(const T* p, Off boff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + boff);
}
template <class T, class Off>
__device__ __forceinline__ void stb(T* p, Off boff, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + boff) = v;
}

template <class T, int MODE, class Off>
__device__ __forceinline__ void nl_point(const NlArgsT<T>& a, int i, int k, Off bo) {
    {
        double u[3], g[3][3];
        const double up0 = (double)ldb(a.ui[0], bo);
        u[0] = up0 + a.ubar[k];
        u[1] = (double)ldb(a.ui[1], bo);
        u[2] = (double)ldb(a.ui[2], bo);
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) g[l][m] = (double)ldb(a.g[l + 3 * m], bo);
        g[0][2] = g[0][2] + a.ubarp[k];
        if (a.write_base) {
            stb(a.uo[0], bo, (T)u[0]);
            stb(a.uo[1], bo, (T)u[1]);
            stb(a.uo[2], bo, (T)u[2]);
            stb(a.go[0 + 3 * 2], bo, (T)g[0][2]);
        }
        // sponge (navierstokes.f90:190-200, 750-756): F = lambda(x) (ubar - u, -v, -w)
        const double lam = a.lam ? a.lam[i] : 0.0;
        const double tgt[3] = {a.ubar[k], 0.0, 0.0};
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) s = s - 0.5 * u[m] * g[l][m];
            if (a.lam) s = s + lam * (tgt[l] - u[l]);
            stb(a.nlt[l], bo, (T)s);
        }
        double tau[3][3];
        if (MODE == NL_SIGMA) {
            const double nu = sigma_nu(g, a.delta[k]);
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m) tau[l][m] = 2.0 * nu * 0.5 * (g[l][m] + g[m][l]);
        } else if (MODE == NL_STATIC) {
            double sn = 0.0;
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) {
                    const double sij = 0.5 * (g[ii][jj] + g[jj][ii]);
                    sn = sn + sij * sij;
                }
            sn = sqrt(2. * sn);
            const double nu = a.smag[k] * sn;
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) tau[ii][jj] = nu * 0.5 * (g[ii][jj] + g[jj][ii]);
        } else if (MODE == NL_EXTTAU) {
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m) tau[l][m] = a.tau[l + 3 * m] ? (double)ldb(a.tau[l + 3 * m], bo) : 0.0;
        } else {
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m) tau[l][m] = 0.0;
        }
        if (MODE == NL_PREP) {
            // products only (dynamic model computes tau from them next)
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    if (a.h[l + 3 * m]) stb(a.h[l + 3 * m], bo, (T)(u[l] * u[m]));
        } else {
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    if (a.h[l + 3 * m]) {
                        double hv = 0.5 * (u[l] * u[m]) + tau[l][m];
                        // u0^2 - ubar^2 = up0 (u0 + ubar): the perturbation part without cancellation
                        if (l == 0 && m == 0 && a.sub_base) hv = 0.5 * (up0 * (u[0] + a.ubar[k])) + tau[0][0];
                        stb(a.h[l + 3 * m], bo, (T)hv);
                    }
        }
    }
}

// One x-row per block row: blockIdx.y strides over the rows k*ny + j, so k and the row
// offset are uniform (scalar) and each thread computes one 32-bit byte offset.
template <class T, int MODE>
__global__ void __launch_bounds__(256) k_nonlinear(PhDims d, NlArgsT<T> a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx) return;
    const int nrows = d.ny * d.nzp;
    for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
        const int k = row / d.ny;
        const unsigned bo = ((unsigned)row * (unsigned)d.nxpp + (unsigned)i) * (unsigned)sizeof(T);
        nl_point<T, MODE>(a, i, k, bo);
    }
}

// the same for fields of 4 GiB and more (64-bit offsets)
template <class T, int MODE>
__global__ void __launch_bounds__(256) k_nonlinear_big(PhDims d, NlArgsT<T> a) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        nl_point<T, MODE>(a, i, k, (size_t)off * sizeof(T));
    }
}

// h_lm = 1/2 uu_lm + tau_lm (after the dynamic model)
__global__ void k_hcombine(PhDims d, FieldList<double> uu, FieldList<double> tau, FieldList<double> h, int n) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        for (int q = 0; q < n; ++q) h.p[q][off] = 0.5 * uu.p[q][off] + tau.p[q][off];
    }
}

// sgsterm.f90:59-81 utau: sum over the wall planes of xnu*du/dz, then the
// per-plane factor -(CS delta (1 - exp(-z+/25)))^2 (sgsterm.f90:84-108).
template <class T>
__global__ void k_utau_reduce(const T* __restrict__ g02, const double* __restrict__ ubarp, int nx, int ny,
                              int nxpp, const int* __restrict__ walls, int nwall, double xnu,
                              double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    const long n = (long)nx * ny;
    for (int w = 0; w < nwall; ++w) {
        const int k = walls[w];
        for (long t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = (int)(t % nx), j = (int)(t / nx);
            const double v = (double)g02[((long)k * ny + j) * nxpp + i] + (ubarp ? ubarp[k] : 0.0);
            s += xnu * v;
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}
__global__ void k_smag_fac(const double* __restrict__ sum, int nwall, long nxny, const double* __restrict__ zpts,
                           const double* __restrict__ delta, int nzp, double xnu, double cs,
                           double* __restrict__ fac) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nzp) return;
    const double utau_avg = sqrt(fabs(sum[0] / nwall / (double)nxny));
    const double zplus = zpts[k] * utau_avg / xnu;
    const double v = cs * delta[k] * (1. - exp(-zplus / 25.));
    fac[k] = -(v * v);
}

// -------------------------------------------------- dynamic Smagorinsky
// stage 1 (sgsterm.f90:171-205): S_ij, |S|, L1_ij = delta^2 |S| S_ij
__global__ void k_dyn1(PhDims d, DynArgs a) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        double g[3][3];
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) g[l][m] = a.g[l + 3 * m][off];
        double sij[3][3];
        double sn = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                sij[l][m] = 0.5 * (g[l][m] + g[m][l]);
                sn = sn + 2. * sij[l][m] * sij[l][m];
            }
        sn = sqrt(sn);
        a.snorm[off] = sn;
        const double dl = a.delta[k];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                const int q = l + 3 * m;
                a.sij[q][off] = sij[l][m];
                a.l1[q][off] = dl * dl * sn * sij[l][m];
            }
    }
}
// stage 2 (sgsterm.f90:187-238): |S^|, M_ij, L_ij, numerator and denominator
__global__ void k_dyn2(PhDims d, DynArgs a) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        double sh[3][3];
        double snh = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                sh[l][m] = a.sijhat[l + 3 * m][off];
                snh = snh + 2. * sh[l][m] * sh[l][m];
            }
        snh = sqrt(snh);
        const double cd = 2.0 * a.delta[k];
        double uh[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) uh[l] = a.uihat[l][off];
        double num = 0.0, den = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                const int q = l + 3 * m;
                const double shat = cd * cd * snh * sh[l][m];
                const double mij = 2.0 * (a.l1hat[q][off] - shat);
                const double lij = a.uuhat[q][off] - uh[l] * uh[m];
                num = num + mij * lij;
                den = den + mij * mij;
            }
        a.num[off] = num;
        a.den[off] = den;
    }
}
// stage 3 (sgsterm.f90:248-267): C = <ML>/<MM> clipped to [0, 0.04], tau = -2 C delta^2 |S| S_ij;
// <.> the local filter (numf, denf) or the plane means (cplane).  With h set, h = uu/2 + tau is
// written over uu instead of tau (the step's h_lm = 1/2 u_l u_m + tau_lm, navierstokes.f90:769),
// once per distinct component.
__global__ void k_dyn3(PhDims d, DynArgs a) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        double ct = a.cplane ? a.cplane[k] / a.cplane[d.nzp + k] : a.numf[off] / a.denf[off];
        if (ct < 0.0 || isnan(ct)) ct = 0.0;
        if (ct > 0.2 * 0.2) ct = 0.2 * 0.2;
        const double dl = a.delta[k];
        const double nu = ct * dl * dl * a.snorm[off];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                const int q = l + 3 * m;
                const double tau = -2.0 * nu * a.sij[q][off];
                if (a.h[q]) {
                    if (a.hsym && l < m) continue;  // (l, m) and (m, l) share one field
                    a.h[q][off] = 0.5 * a.uu[q][off] + tau;
                } else if (a.tau[q]) {
                    a.tau[q][off] = tau;
                }
            }
    }
}

// ----------------------------------------------------- interscale model
// sgsterm.f90:448-462: uibuj_lm = [u_l] u_m and tau = [[u_l]u_m] - [[u_l]][[u_m]] - [u_l](u_m - [u_m])
__global__ void k_isc1(PhDims d, FieldList<double> ub, FieldList<double> ui, FieldList<double> out) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l) out.p[l + 3 * m][off] = ub.p[l][off] * ui.p[m][off];
    }
}
__global__ void k_isc2(PhDims d, FieldList<double> ubb, FieldList<double> ub2, FieldList<double> ub,
                       FieldList<double> ui, FieldList<double> tau) {
    PH_LOOP(d) {
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int l = 0; l < 3; ++l)
                tau.p[l + 3 * m][off] = ubb.p[l + 3 * m][off] - ub2.p[l][off] * ub2.p[m][off] -
                                        ub.p[l][off] * (ui.p[m][off] - ub.p[m][off]);
    }
}

// ------------------------------------------------------------- 3-pt filter
// filters.f90:363-613 at one rank: x then y periodic (a,b,a)/(2a+b) passes
// (fused, rounding after each pass as in the reference), then the z pass with
// interpolating weights on interior planes (k = 0 and nzp-1 keep the xy value).
__global__ void k_filt_xy(PhDims d, FieldList<double> in, FieldList<double> out, int nf, double a, double b) {
    const long n = d.npts * nf;
    const double s = 2. * a + b;
    const double s2 = 2.00 * a + b;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int f = (int)(t / d.npts);
        const long e = t - (long)f * d.npts;
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        const double* q = in.p[f];
        const int im = (i == 0) ? d.nx - 1 : i - 1;
        const int ip = (i == d.nx - 1) ? 0 : i + 1;
        const int jm = (j == 0) ? d.ny - 1 : j - 1;
        const int jp = (j == d.ny - 1) ? 0 : j + 1;
        const long pk = (long)k * d.ny;
        double X[3];
        const int js[3] = {jm, j, jp};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double* row = q + (pk + js[r]) * d.nxpp;
            X[r] = (a * row[im] + b * row[i] + a * row[ip]) / s;
        }
        out.p[f][off] = (a * X[0] + b * X[1] + a * X[2]) / s2;
    }
}
__global__ void k_filt_z(PhDims d, FieldList<double> in, FieldList<double> out, int nf, const double* __restrict__ w,
                         double a, double b) {
    const long n = d.npts * nf;
    const double s = 2.0 * a + b;
    const long plane = (long)d.ny * d.nxpp;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int f = (int)(t / d.npts);
        const long e = t - (long)f * d.npts;
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        const double* q = in.p[f];
        double v;
        if (k == 0 || k == d.nzp - 1) {
            v = q[off];
        } else {
            v = w[k] * q[off - plane] + w[d.nzp + k] * q[off] + w[2 * d.nzp + k] * q[off + plane];
            v = v / s;
        }
        out.p[f][off] = v;
    }
}


// The same filter in one pass: block (R rows j, KC planes k, field) slides over k with the
// xy-filtered values of planes k-1, k, k+1 in registers (threads over x, NXT points each), so
// every input value is read from HBM about once (halo rows and planes aside) and the
// xy-filtered field never goes through HBM.  Same per-element expressions as the two passes.
// T is the storage type (double; float in the mixed form); the arithmetic is FP64.
template <class T, int R, int NXT>
__global__ void __launch_bounds__(256) k_filt3_fused(PhDims d, FieldList<T> in, FieldList<T> out, double a,
                                                     double b, const double* __restrict__ w, int KC) {
    const int f = blockIdx.z;
    const int J0 = blockIdx.x * R;
    const int k0 = blockIdx.y * KC;
    const int k1 = min(k0 + KC, d.nzp);
    const T* q = in.p[f];
    T* o = out.p[f];
    const double s = 2. * a + b;
    const double s2 = 2.00 * a + b;
    const bool p2 = pow2_exact(s);
    const double is = 1.0 / s;
    const long plane = (long)d.ny * d.nxpp;
    int jr[R + 2];
#pragma unroll
    for (int rr = 0; rr < R + 2; ++rr) {
        int j = J0 - 1 + rr;
        j = (j < 0) ? j + d.ny : j;
        jr[rr] = j % d.ny;
    }
    auto xy = [&](int k, double (&y)[R][NXT]) {
        double X[R + 2][NXT];
#pragma unroll
        for (int rr = 0; rr < R + 2; ++rr) {
            const T* row = q + (long)k * plane + (long)jr[rr] * d.nxpp;
#pragma unroll
            for (int t = 0; t < NXT; ++t) {
                const int i = threadIdx.x + 256 * t;
                if (i < d.nx) {
                    const int im = (i == 0) ? d.nx - 1 : i - 1;
                    const int ip = (i == d.nx - 1) ? 0 : i + 1;
                    X[rr][t] = fdiv(a * (double)row[im] + b * (double)row[i] + a * (double)row[ip], s, is, p2);
                } else {
                    X[rr][t] = 0.0;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int t = 0; t < NXT; ++t) y[r][t] = fdiv(a * X[r][t] + b * X[r + 1][t] + a * X[r + 2][t], s2, is, p2);
    };
    double Yp[R][NXT], Yc[R][NXT], Yn[R][NXT];
    xy(k0, Yc);
    if (k0 > 0) xy(k0 - 1, Yp);
    for (int k = k0; k < k1; ++k) {
        if (k + 1 < d.nzp) xy(k + 1, Yn);
        const bool interior = (k > 0 && k < d.nzp - 1);
        const double w0 = w[k], w1 = w[d.nzp + k], w2 = w[2 * d.nzp + k];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = J0 + r;
            if (j >= d.ny) continue;
            T* orow = o + (long)k * plane + (long)j * d.nxpp;
#pragma unroll
            for (int t = 0; t < NXT; ++t) {
                const int i = threadIdx.x + 256 * t;
                if (i >= d.nx) continue;
                double v;
                if (interior) {
                    v = w0 * Yp[r][t] + w1 * Yc[r][t] + w2 * Yn[r][t];
                    v = fdiv(v, s, is, p2);
                } else {
                    v = Yc[r][t];
                }
                orow[i] = (T)v;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int t = 0; t < NXT; ++t) {
                Yp[r][t] = Yc[r][t];
                Yc[r][t] = Yn[r][t];
            }
    }
}

// The same filter through LDS: a block owns an x-tile of TX points by R rows of j and
// slides over KC planes of k.  Each plane's (R + 2) x (TX + 2) input tile (periodic halos)
// goes global -> registers -> LDS, the loads of the plane after next in flight while the
// next plane is filtered; the x pass fills a second LDS array, the y pass gives the
// xy-filtered value of the thread's points, and the z pass combines the three planes held
// in registers.  Same per-element expressions and order as the two-pass kernels, so the
// results are bit-identical to them.
template <class T, int R, int TX = 64>
__global__ void __launch_bounds__(256) k_filt3_lds(PhDims d, FieldList<T> in, FieldList<T> out, double a, double b,
                                                   const double* __restrict__ w, int KC) {
    constexpr int NT = (R + 2) * (TX + 2);          // tile elements
    constexpr int NL = (NT + 255) / 256;            // tile loads per thread
    constexpr int NP = (R * TX) / 256;              // output points per thread
    static_assert((R * TX) % 256 == 0, "R * TX must be a multiple of 256");
    __shared__ T tile[NT];
    __shared__ double xs[(R + 2) * TX];
    const int f = blockIdx.z;
    const int ntx = (d.nx + TX - 1) / TX;
    const int i0 = (blockIdx.x % ntx) * TX;
    const int J0 = (blockIdx.x / ntx) * R;
    const int k0 = blockIdx.y * KC;
    const int k1 = min(k0 + KC, d.nzp);
    const T* q = in.p[f];
    T* o = out.p[f];
    const double s = 2. * a + b;
    const double s2 = 2.00 * a + b;
    const bool p2 = pow2_exact(s);
    const double is = 1.0 / s;
    const long plane = (long)d.ny * d.nxpp;
    // this thread's tile elements (offsets within a plane; -1 past the tile)
    int toff[NL];
#pragma unroll
    for (int m = 0; m < NL; ++m) {
        const int e = threadIdx.x + 256 * m;
        int off = -1;
        if (e < NT) {
            const int rr = e / (TX + 2), cc = e - rr * (TX + 2);
            int j = (J0 - 1 + rr) % d.ny;
            j = (j < 0) ? j + d.ny : j;
            int i = (i0 - 1 + cc) % d.nx;
            i = (i < 0) ? i + d.nx : i;
            off = j * d.nxpp + i;
        }
        toff[m] = off;
    }
    T pre[NL];
    auto load = [&](int k) {
        const T* qk = q + (long)k * plane;
#pragma unroll
        for (int m = 0; m < NL; ++m) pre[m] = (toff[m] >= 0) ? qk[toff[m]] : (T)0;
    };
    // xy-filtered values (y) of the plane whose tile is in `pre`; then the loads of plane
    // knext (>= 0) into `pre`, in flight during the x and y passes
    auto stage = [&](double (&y)[NP], int knext) {
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int e = threadIdx.x + 256 * m;
            if (e < NT) tile[e] = pre[m];
        }
        __syncthreads();
        if (knext >= 0) load(knext);
        for (int e = threadIdx.x; e < (R + 2) * TX; e += 256) {
            const int rr = e / TX, c = e - rr * TX;
            const T* t = tile + rr * (TX + 2) + c;
            xs[e] = fdiv(a * (double)t[0] + b * (double)t[1] + a * (double)t[2], s, is, p2);
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const int p = threadIdx.x + 256 * m;
            const int r = p / TX, c = p - r * TX;
            y[m] = fdiv(a * xs[r * TX + c] + b * xs[(r + 1) * TX + c] + a * xs[(r + 2) * TX + c], s2, is, p2);
        }
    };
    // planes needed: k0 - 1 (if any) .. min(k1, nzp - 1)
    const int klast = min(k1, d.nzp - 1);
    double Yp[NP], Yc[NP], Yn[NP];
    if (k0 > 0) {
        load(k0 - 1);
        stage(Yp, k0);
    } else {
        load(k0);
    }
    stage(Yc, (k0 + 1 <= klast) ? k0 + 1 : -1);
    for (int k = k0; k < k1; ++k) {
        if (k + 1 <= klast) stage(Yn, (k + 2 <= klast) ? k + 2 : -1);
        const bool interior = (k > 0 && k < d.nzp - 1);
        const double w0 = w[k], w1 = w[d.nzp + k], w2 = w[2 * d.nzp + k];
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const int p = threadIdx.x + 256 * m;
            const int r = p / TX, c = p - r * TX;
            const int j = J0 + r, i = i0 + c;
            if (j < d.ny && i < d.nx) {
                double v;
                if (interior) {
                    v = w0 * Yp[m] + w1 * Yc[m] + w2 * Yn[m];
                    v = fdiv(v, s, is, p2);
                } else {
                    v = Yc[m];
                }
                o[(long)k * plane + (long)j * d.nxpp + i] = (T)v;
            }
            Yp[m] = Yc[m];
            Yc[m] = Yn[m];
        }
    }
}

template <class T, int R, int TX = 64>
static void launch_filt3_lds(const PhDims& d, const FieldList<T>& fi, const FieldList<T>& fo, int nb, double a,
                             double b, const double* w, int KC, hipStream_t s) {
    const int ntx = (d.nx + TX - 1) / TX;
    const dim3 grid((unsigned)(ntx * ((d.ny + R - 1) / R)), (unsigned)((d.nzp + KC - 1) / KC), (unsigned)nb);
    hipLaunchKernelGGL((k_filt3_lds<T, R, TX>), grid, dim3(256), 0, s, d, fi, fo, a, b, w, KC);
}

template <class T, int R, int NXT>
static void launch_filt3_fused(const PhDims& d, const FieldList<T>& fi, const FieldList<T>& fo, int nb, double a,
                               double b, const double* w, int KC, hipStream_t s) {
    const dim3 grid((unsigned)((d.ny + R - 1) / R), (unsigned)((d.nzp + KC - 1) / KC), (unsigned)nb);
    hipLaunchKernelGGL((k_filt3_fused<T, R, NXT>), grid, dim3(256), 0, s, d, fi, fo, a, b, w, KC);
}

// Launch shapes of the fused filter: R rows of j per block, KC planes of k per block
// (the x extent is the whole row, NXT points per thread).  The variant comes from the
// context's tuning table (lsb_autotune), or the default.
static const int kFiltR[3] = {2, 4, 8};
static const int kFiltKC[3] = {8, 16, 32};
// variants 0..8: the register-sliding kernel (R, KC) above; 9..17: the LDS kernel with
// 64-point x tiles, R in {4, 8, 16} and KC in {16, 32, 64}; 18..23: 128-point x tiles
// (twice the outputs per thread and barrier, half the relative x halo), R in {4, 8}
static const int kLdsR[3] = {4, 8, 16};
static const int kLdsKC[3] = {16, 32, 64};
static const int kLdsWideR[2] = {4, 8};
constexpr int kFilt3Variants = 24;
// defaults (measured on 512 x 193 x 256, 8 fields per launch): LDS kernel, FP64 storage
// 64-point tiles, R = 4, KC = 16 (0.77 ms); FP32 storage 128-point tiles, R = 4, KC = 16
// (0.43 ms against 0.46 with the best 64-point tile)
constexpr int kFilt3Default64 = 9;
constexpr int kFilt3Default32 = 18;

template <class T, int NXT>
static void filt3_launch_r(const PhDims& d, const FieldList<T>& fi, const FieldList<T>& fo, int nb, double a,
                           double b, const double* w, int R, int KC, hipStream_t s) {
    switch (R) {
        case 2: launch_filt3_fused<T, 2, NXT>(d, fi, fo, nb, a, b, w, KC, s); break;
        case 4: launch_filt3_fused<T, 4, NXT>(d, fi, fo, nb, a, b, w, KC, s); break;
        default: launch_filt3_fused<T, 8, NXT>(d, fi, fo, nb, a, b, w, KC, s); break;
    }
}

template <class T>
static int filt3_variant(const PhDims& d, const FieldList<T>& fi, const FieldList<T>& fo, int nb, double a, double b,
                         const double* w, int v, hipStream_t s) {
    if (v >= 18) {
        const int R = kLdsWideR[(v - 18) / 3], KC = kLdsKC[(v - 18) % 3];
        if (R == 4)
            launch_filt3_lds<T, 4, 128>(d, fi, fo, nb, a, b, w, KC, s);
        else
            launch_filt3_lds<T, 8, 128>(d, fi, fo, nb, a, b, w, KC, s);
        LSB_HIP(hipGetLastError());
        return LSB_OK;
    }
    if (v >= 9) {
        const int R = kLdsR[(v - 9) / 3], KC = kLdsKC[(v - 9) % 3];
        switch (R) {
            case 4: launch_filt3_lds<T, 4>(d, fi, fo, nb, a, b, w, KC, s); break;
            case 8: launch_filt3_lds<T, 8>(d, fi, fo, nb, a, b, w, KC, s); break;
            default: launch_filt3_lds<T, 16>(d, fi, fo, nb, a, b, w, KC, s); break;
        }
        LSB_HIP(hipGetLastError());
        return LSB_OK;
    }
    const int R = kFiltR[v / 3], KC = kFiltKC[v % 3];
    if (d.nx <= 256)
        filt3_launch_r<T, 1>(d, fi, fo, nb, a, b, w, R, KC, s);
    else if (d.nx <= 512)
        filt3_launch_r<T, 2>(d, fi, fo, nb, a, b, w, R, KC, s);
    else
        filt3_launch_r<T, 4>(d, fi, fo, nb, a, b, w, R, KC, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

template <class T>
static int filt3_dispatch(lsb_ctx* c, const PhDims& d, const FieldList<T>& fi, const FieldList<T>& fo, int nb,
                          double a, double b, const double* w, hipStream_t s) {
    int v = c->tune[sizeof(T) == 4 ? TUNE_FILT3_F32 : TUNE_FILT3_F64];
    if (v < 0 || v >= kFilt3Variants) v = (sizeof(T) == 4) ? kFilt3Default32 : kFilt3Default64;
    return filt3_variant<T>(d, fi, fo, nb, a, b, w, v, s);
}

// ------------------------------------------------------ mode-cut accumulation
// filters.f90:640-646: qb = qb + coef*q_filt over the physical points (first pass: coef*q).
__global__ void k_axpy_fields(PhDims d, FieldList<const double> q, FieldList<double> qb, int nf, double coef,
                              int first) {
    const long n = d.npts * nf;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int f = (int)(t / d.npts);
        const long e = t - (long)f * d.npts;
        int i, j, k;
        long off;
        ph_index(d, e, i, j, k, off);
        const double v = coef * q.p[f][off];
        qb.p[f][off] = first ? v : qb.p[f][off] + v;
    }
}

// filters.f90:103-108: per plane k, sum over the nf fields and the nx*ny physical points of
// (q - f)^2.  Block (k, part): part of the rows of plane k; one partial per block,
// part[k*kEnergyParts + p], added on the host in a fixed order.
constexpr int kEnergyParts = 8;
__global__ void __launch_bounds__(256) k_plane_energy(PhDims d, FieldList<const double> q, FieldList<const double> f,
                                                      int nf, double* __restrict__ part) {
    const int k = blockIdx.x, p = blockIdx.y;
    const long per_plane = (long)d.ny * d.nx;
    double s = 0.0;
    for (int l = 0; l < nf; ++l)
        for (long e = (long)p * blockDim.x + threadIdx.x; e < per_plane; e += (long)kEnergyParts * blockDim.x) {
            const int j = (int)(e / d.nx), i = (int)(e - (long)j * d.nx);
            const long off = ((long)k * d.ny + j) * d.nxpp + i;
            const double r = q.p[l][off] - f.p[l][off];
            s += r * r;
        }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(long)k * kEnergyParts + p] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// ------------------------------------------------------------- launchers
int ph_axpy(lsb_ctx* c, const double* const* q, double* const* qb, int nf, double coef, bool first, hipStream_t s) {
    PhDims d = ph_dims(c);
    for (int f0 = 0; f0 < nf; f0 += kMaxFields) {
        const int nb = (nf - f0 < kMaxFields) ? nf - f0 : kMaxFields;
        FieldList<const double> fi{};
        FieldList<double> fo{};
        for (int l = 0; l < nb; ++l) {
            fi.p[l] = q[f0 + l];
            fo.p[l] = qb[f0 + l];
        }
        hipLaunchKernelGGL(k_axpy_fields, ph_grid(d.npts * nb), dim3(256), 0, s, d, fi, fo, nb, coef, first ? 1 : 0);
        LSB_HIP(hipGetLastError());
    }
    return LSB_OK;
}

int ph_plane_energy(lsb_ctx* c, const double* const* q, const double* const* f, int nf, double* e_host,
                    hipStream_t s) {
    if (nf < 1 || nf > kMaxFields || (long)c->nzp * kEnergyParts > kRedDoubles) {
        set_error("plane_energy: bad field count or nzp too large");
        return LSB_E_ARG;
    }
    PhDims d = ph_dims(c);
    FieldList<const double> fq{}, ff{};
    for (int l = 0; l < nf; ++l) {
        fq.p[l] = q[l];
        ff.p[l] = f[l];
    }
    hipLaunchKernelGGL(k_plane_energy, dim3((unsigned)c->nzp, (unsigned)kEnergyParts), dim3(256), 0, s, d, fq, ff, nf,
                       c->d_red);
    LSB_HIP(hipGetLastError());
    std::vector<double> part((size_t)c->nzp * kEnergyParts);
    LSB_HIP(hipMemcpyAsync(part.data(), c->d_red, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    LSB_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < c->nzp; ++k) {
        double v = 0.0;
        for (int p = 0; p < kEnergyParts; ++p) v += part[(size_t)k * kEnergyParts + p];
        e_host[k] = v;
    }
    return LSB_OK;
}

int ph_zero_pads(lsb_ctx* c, double* const* f, int nf, hipStream_t s) {
    FieldList<double> fl{};
    for (int i = 0; i < nf; ++i) fl.p[i] = f[i];
    const long nrows = (long)c->nzp * c->ny;
    hipLaunchKernelGGL(k_zero_pads, ph_grid(nrows * nf * 2), dim3(256), 0, s, fl, nf, nrows, c->nx, c->nxpp);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

template <class T>
static int launch_nonlinear(lsb_ctx* c, int mode, const NlArgsT<T>& a, hipStream_t s) {
    PhDims d = ph_dims(c);
    int tok = -1;
    // 12 fields in, 3 + up to 9 out
    int nh = 0;
    for (int q = 0; q < 9; ++q) nh += a.h[q] ? 1 : 0;
    prof_mark_begin(c, PC_NL, s, &tok, 0.0, (double)(15 + nh) * d.npts * sizeof(T));
    const bool small = (size_t)d.nxpp * d.ny * d.nzp * sizeof(T) < ((size_t)1 << 32);
    const int nrows = d.ny * d.nzp;
    const dim3 g = small ? dim3((unsigned)((d.nx + 255) / 256), (unsigned)std::min(nrows, 65535)) : ph_grid(d.npts);
#define LSB_NL_LAUNCH(M)                                                                   \
    if (small)                                                                             \
        hipLaunchKernelGGL((k_nonlinear<T, M>), g, dim3(256), 0, s, d, a);                 \
    else                                                                                   \
        hipLaunchKernelGGL((k_nonlinear_big<T, M>), g, dim3(256), 0, s, d, a);
    switch (mode) {
        case NL_NONE: LSB_NL_LAUNCH(NL_NONE) break;
        case NL_SIGMA: LSB_NL_LAUNCH(NL_SIGMA) break;
        case NL_STATIC: LSB_NL_LAUNCH(NL_STATIC) break;
        case NL_EXTTAU: LSB_NL_LAUNCH(NL_EXTTAU) break;
        case NL_PREP: LSB_NL_LAUNCH(NL_PREP) break;
        default: set_error("nonlinear: bad mode"); return LSB_E_ARG;
    }
#undef LSB_NL_LAUNCH
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int ph_nonlinear(lsb_ctx* c, int mode, const NlArgs& a, hipStream_t s) { return launch_nonlinear<double>(c, mode, a, s); }
int ph_nonlinear_f32(lsb_ctx* c, int mode, const NlArgsT<float>& a, hipStream_t s) {
    return launch_nonlinear<float>(c, mode, a, s);
}

int ph_hcombine(lsb_ctx* c, const double* const* uu, const double* const* tau, double* const* h, int n, hipStream_t s) {
    PhDims d = ph_dims(c);
    FieldList<double> a{}, b{}, o{};
    for (int q = 0; q < n; ++q) {
        a.p[q] = const_cast<double*>(uu[q]);
        b.p[q] = const_cast<double*>(tau[q]);
        o.p[q] = h[q];
    }
    hipLaunchKernelGGL(k_hcombine, ph_grid(d.npts), dim3(256), 0, s, d, a, b, o, n);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// The wall planes (zpts == 0 or |zpts| == 0.5, sgsterm.f90:59-81) are uploaded once
// with the context (d_walls), so this only enqueues work.
template <class T>
static int smag_factor(lsb_ctx* c, const T* g02, const double* ubarp, double cs, hipStream_t s) {
    if (c->nwalls <= 0) {
        set_error("static_smagorinsky: no wall plane (zpts == 0 or |zpts| == 0.5)");
        return LSB_E_ARG;
    }
    hipLaunchKernelGGL(k_utau_reduce<T>, dim3(1), dim3(256), 0, s, g02, ubarp, c->nx, c->ny, c->nxpp, c->d_walls,
                       c->nwalls, c->cfg.xnu, c->d_red + kRedDoubles - 1);
    LSB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_smag_fac, dim3((c->nzp + 63) / 64), dim3(64), 0, s, c->d_red + kRedDoubles - 1, c->nwalls,
                       (long)c->nx * c->ny, c->d_zpts, c->d_delta, c->nzp, c->cfg.xnu, cs, c->d_smag);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
int ph_smag_factor(lsb_ctx* c, const double* g02, const double* ubarp, double cs, hipStream_t s) {
    return smag_factor<double>(c, g02, ubarp, cs, s);
}
int ph_smag_factor_f32(lsb_ctx* c, const float* g02, const double* ubarp, double cs, hipStream_t s) {
    return smag_factor<float>(c, g02, ubarp, cs, s);
}

// The z weights on the device for the pairs the step uses, or null.
static const double* filt_weights(const lsb_ctx* c, double a, double b) {
    if (a == c->cfg.a && b == c->cfg.b) return c->d_fz;
    if (a == 1.0 && b == 1.0) return c->d_fz + 3 * c->nzp;
    if (a == 1.0 && b == 6.0) return c->d_fz + 6 * c->nzp;
    if (a == 1.0 && b == 2.0) return c->d_fz + 9 * c->nzp;
    return nullptr;
}

// LSB_FILT3_TWO_PASS=1 selects the two-pass kernels (x-y, then z through tmp)
static bool filt3_two_pass() {
    static const bool v = [] {
        const char* e = std::getenv("LSB_FILT3_TWO_PASS");
        return e && e[0] == '1';
    }();
    return v;
}

int ph_filter3(lsb_ctx* c, const double* const* in, double* const* out, int nf, double a, double b, int interpflag,
               double* tmp, hipStream_t s) {
    PhDims d = ph_dims(c);
    int tok = -1;
    // x-y pass and z pass, each one read and one write per field
    prof_mark_begin(c, PC_PW, s, &tok, 0.0, 4.0 * nf * (double)d.npts * sizeof(double));
    const double* w = filt_weights(c, a, b);
    std::vector<double> hw;
    double* wtmp = nullptr;
    if (interpflag != 1 || w == nullptr) {
        // general (a, b) or interpflag = 0: uniform weights a, b, a or interpolating with (a, b)
        hw.assign(3 * c->nzp, 0.0);
        if (interpflag != 1) {
            for (int k = 0; k < c->nzp; ++k) {
                hw[k] = a;
                hw[c->nzp + k] = b;
                hw[2 * c->nzp + k] = a;
            }
        } else {
            set_error("filter_q_3pt: (a, b) must be the configured pair, (1, 1), (1, 6) or (1, 2)");
            return LSB_E_ARG;
        }
        wtmp = c->d_red + 64;
        LSB_HIP(hipMemcpyAsync(wtmp, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice, s));
        w = wtmp;
    }
    for (int f0 = 0; f0 < nf; f0 += kMaxFields) {
        const int nb = (nf - f0 < kMaxFields) ? nf - f0 : kMaxFields;
        FieldList<double> fi{}, ft{}, fo{};
        for (int q = 0; q < nb; ++q) {
            fi.p[q] = const_cast<double*>(in[f0 + q]);
            ft.p[q] = tmp + (size_t)q * c->nphys;
            fo.p[q] = out[f0 + q];
        }
        if (d.nx <= 1024 && !filt3_two_pass()) {
            LSB_TRY(filt3_dispatch<double>(c, d, fi, fo, nb, a, b, w, s));
        } else {
            hipLaunchKernelGGL(k_filt_xy, ph_grid(d.npts * nb), dim3(256), 0, s, d, fi, ft, nb, a, b);
            LSB_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_filt_z, ph_grid(d.npts * nb), dim3(256), 0, s, d, ft, fo, nb, w, a, b);
        }
        LSB_HIP(hipGetLastError());
    }
    prof_mark_end(c, tok, s);
    if (wtmp) LSB_HIP(hipStreamSynchronize(s));
    return LSB_OK;
}

// FP32 storage, FP64 arithmetic (the mixed form's dynamic model); interpolating z weights.
int ph_filter3_f32(lsb_ctx* c, const float* const* in, float* const* out, int nf, double a, double b, hipStream_t s) {
    PhDims d = ph_dims(c);
    const double* w = filt_weights(c, a, b);
    if (!w || d.nx > 1024) {
        set_error("filter (FP32): needs nx <= 1024 and the configured (a, b) or (1, 1)");
        return LSB_E_UNSUPPORTED;
    }
    int tok = -1;
    prof_mark_begin(c, PC_PW, s, &tok, 0.0, 2.0 * nf * (double)d.npts * sizeof(float));
    for (int f0 = 0; f0 < nf; f0 += kMaxFields) {
        const int nb = (nf - f0 < kMaxFields) ? nf - f0 : kMaxFields;
        FieldList<float> fi{}, fo{};
        for (int q = 0; q < nb; ++q) {
            fi.p[q] = const_cast<float*>(in[f0 + q]);
            fo.p[q] = out[f0 + q];
        }
        LSB_TRY(filt3_dispatch<float>(c, d, fi, fo, nb, a, b, w, s));
    }
    prof_mark_end(c, tok, s);
    return LSB_OK;
}

// Times every launch variant of the fused filter on nf scratch fields of storage type T
// (the first 2 nf physical fields of `work`) and records the fastest in the context.
template <class T>
static int tune_filt3_t(lsb_ctx* c, void* work, int nf, hipStream_t s) {
    PhDims d = ph_dims(c);
    if (d.nx > 1024) return LSB_OK;
    const double* w = c->d_fz;
    FieldList<T> fi{}, fo{};
    for (int q = 0; q < nf; ++q) {
        fi.p[q] = (T*)work + (size_t)q * c->nphys;
        fo.p[q] = (T*)work + (size_t)(nf + q) * c->nphys;
    }
    hipEvent_t e0, e1;
    LSB_HIP(hipEventCreate(&e0));
    LSB_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bv = (sizeof(T) == 4) ? kFilt3Default32 : kFilt3Default64;
    for (int v = 0; v < kFilt3Variants; ++v) {
        LSB_TRY(filt3_variant<T>(d, fi, fo, nf, c->cfg.a, c->cfg.b, w, v, s));  // warm-up
        LSB_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) LSB_TRY(filt3_variant<T>(d, fi, fo, nf, c->cfg.a, c->cfg.b, w, v, s));
        LSB_HIP(hipEventRecord(e1, s));
        LSB_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LSB_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (std::getenv("LSB_TUNE_VERBOSE"))
            std::fprintf(stderr, "lsb_autotune filter3 %s variant %d: %.4f ms per launch of %d fields\n",
                         sizeof(T) == 4 ? "f32" : "f64", v, ms / 3.0f, nf);
        if (ms < best) {
            best = ms;
            bv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int slot = sizeof(T) == 4 ? TUNE_FILT3_F32 : TUNE_FILT3_F64;
    c->tune[slot] = bv;
    c->tune_ms[slot] = best / 3.0f;
    return LSB_OK;
}

int ph_tune_filt3(lsb_ctx* c, void* work, int nf, hipStream_t s) {
    LSB_TRY(tune_filt3_t<double>(c, work, nf, s));
    return tune_filt3_t<float>(c, work, nf, s);
}

int ph_dyn(lsb_ctx* c, int stage, const DynArgs& a, hipStream_t s) {
    PhDims d = ph_dims(c);
    switch (stage) {
        case 1: hipLaunchKernelGGL(k_dyn1, ph_grid(d.npts), dim3(256), 0, s, d, a); break;
        case 2: hipLaunchKernelGGL(k_dyn2, ph_grid(d.npts), dim3(256), 0, s, d, a); break;
        default: hipLaunchKernelGGL(k_dyn3, ph_grid(d.npts), dim3(256), 0, s, d, a); break;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int ph_interscale(lsb_ctx* c, int stage, const FieldList<double>& p0, const FieldList<double>& p1,
                  const FieldList<double>& p2, const FieldList<double>& p3, const FieldList<double>& p4,
                  hipStream_t s) {
    PhDims d = ph_dims(c);
    if (stage == 1)
        hipLaunchKernelGGL(k_isc1, ph_grid(d.npts), dim3(256), 0, s, d, p0, p1, p2);
    else
        hipLaunchKernelGGL(k_isc2, ph_grid(d.npts), dim3(256), 0, s, d, p0, p1, p2, p3, p4);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

}  // namespace lsb
