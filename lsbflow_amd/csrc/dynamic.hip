// Dynamic Smagorinsky model of the mixed form (sgsterm.f90:123-280) and the plane means of
// the plane-averaged coefficient (both forms).
//
// Mixed form: every work field is stored in FP32 and every expression is evaluated in FP64
// registers.  The fields hold perturbation quantities only; the base flow U = (ubar(z), 0, 0)
// enters the kernels as FP64 profiles, so no FP32 field carries an O(1) part that later
// cancels:
//   stage 1  from u', g' (perturbation velocity and gradient): the advective term, the
//            perturbation products p_lm = u_l u_m - U_l U_m, the perturbation strain
//            S'_lm = (g'_lm + g'_ml)/2, |S| and L1_lm = delta^2 |S| S_lm of the full strain;
//   filter   S', L1, u', p with the (a, b) test filter;
//   stage 2  M_lm = 2 (L1^ - (2 delta)^2 |S^| S^_lm) with S^ = S'^ + (S^U)^ and
//            L_lm = (u_l u_m)^ - u^_l u^_m = p^_lm - (U^_l u'^_m + u'^_l U^_m + u'^_l u'^_m)
//                   + [(U^2)^ - (U^)^2] (0,0 only),
//            numerator M:L and denominator M:M;
//   <.>      (1, 1) filter of both, or their plane means;
//   stage 3  C = <M:L>/<M:M> clipped to [0, 0.04], tau = -2 C delta^2 |S| S, and
//            h_lm = p_lm/2 + tau_lm written over p (h_00 without its plane-constant ubar^2/2,
//            as the mixed form's other models store it).
// The z-filtered base profiles U^, (S^U)^_02 and (U^2)^ - (U^)^2 are precomputed in FP64.
// Symmetric tensors are stored once per distinct component in the order
// (0,0), (1,0), (2,0), (1,1), (2,1), (2,2).

#include <cmath>
#include <cstdlib>

#include "lsb_internal.h"
#include "physical.h"

namespace lsb {

namespace {

struct DDims {
    int nx, ny, nxpp, nzp;
};

DDims ddims(const lsb_ctx* c) { return DDims{c->nx, c->ny, c->nxpp, c->nzp}; }

constexpr int kL[6] = {0, 1, 2, 1, 2, 2};
constexpr int kM[6] = {0, 0, 0, 1, 1, 2};

// grid: x in blocks of 256 points, one (j, k) row per blockIdx.y (strided)
dim3 row_grid(const DDims& d) {
    const unsigned gx = (unsigned)((d.nx + 255) / 256);
    const long rows = (long)d.ny * d.nzp;
    return dim3(gx, (unsigned)(rows < 65535 ? rows : 65535));
}

// FP32 field element by byte offset: with a 32-bit offset the loads and stores take the
// scalar-base + vector-offset form (no 64-bit address arithmetic per field)
template <class Off>
__device__ __forceinline__ float ldf(const float* p, Off bo) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + bo);
}
template <class Off>
__device__ __forceinline__ void stf(float* p, Off bo, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + bo) = v;
}

}  // namespace

// stage 1 at one point of plane k (column i) from u' and g'; out: nlt (3), p (6), |S|,
// S' (6), L1 (6) in that order
__device__ __forceinline__ void dynm_stage1_math(const DynMixArgs& a, int k, int i, const double (&up)[3],
                                                 const double (&g)[3][3], float (&out)[22]) {
    const double U = a.ubar[k];
    double u[3];
    u[0] = up[0] + U;
    u[1] = up[1];
    u[2] = up[2];
    // advective term -1/2 u_m du_l/dx_m with the full gradient (g'_02 + ubarp), sponge
    const double g02 = g[0][2] + a.ubarp[k];
    const double lam = a.lam ? a.lam[i] : 0.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) s = s - 0.5 * u[m] * ((l == 0 && m == 2) ? g02 : g[l][m]);
        if (a.lam) s = s - lam * up[l];  // lambda (target - u): target = ubar for u, 0 for v, w
        out[l] = (float)s;
    }
    // perturbation products: u_l u_m - U_l U_m (only U_0 = ubar is nonzero)
    out[3] = (float)(up[0] * (u[0] + U));
    out[4] = (float)(up[1] * u[0]);
    out[5] = (float)(up[2] * u[0]);
    out[6] = (float)(up[1] * up[1]);
    out[7] = (float)(up[2] * up[1]);
    out[8] = (float)(up[2] * up[2]);
    // strain of the perturbation and of the full field
    double sp[6], sn = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) sp[q] = 0.5 * (g[kL[q]][kM[q]] + g[kM[q]][kL[q]]);
    const double su = 0.5 * a.ubarp[k];  // base strain S^U_20 = S^U_02
    double sf[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        sf[q] = sp[q] + (q == 2 ? su : 0.0);
        const double w = (kL[q] == kM[q]) ? 1.0 : 2.0;  // off-diagonals appear twice
        sn = sn + w * 2.0 * sf[q] * sf[q];
    }
    sn = sqrt(sn);
    const double dl = a.delta[k];
    out[9] = (float)sn;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        out[10 + q] = (float)sp[q];
        out[16 + q] = (float)(dl * dl * sn * sf[q]);
    }
}

__device__ __forceinline__ float* dynm_stage1_dst(const DynMixArgs& a, int n) {
    return n < 3 ? a.nlt[n] : n < 9 ? a.p[n - 3] : n == 9 ? a.snorm : n < 16 ? a.sp[n - 10] : a.l1[n - 16];
}

template <class Off>
__global__ void __launch_bounds__(256) k_dynm_stage1(DDims d, DynMixArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx) return;
    const int nrows = d.ny * d.nzp;
    for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
        const int k = row / d.ny;
        const Off o = ((Off)row * (Off)d.nxpp + (Off)i) * (Off)sizeof(float);
        double up[3], g[3][3];
#pragma unroll
        for (int l = 0; l < 3; ++l) up[l] = (double)ldf(a.ui[l], o);
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) g[l][m] = (double)ldf(a.g[l + 3 * m], o);
        float out[22];
        dynm_stage1_math(a, k, i, up, g, out);
#pragma unroll
        for (int n = 0; n < 22; ++n) stf(dynm_stage1_dst(a, n), o, out[n]);
    }
}

// The same with two adjacent x points per thread: 8-byte loads and stores (nx and nxpp
// even, so every pair is 8-byte aligned), the per-point expressions unchanged.
__global__ void __launch_bounds__(256) k_dynm_stage1x2(DDims d, DynMixArgs a) {
    const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.nx) return;
    const int nrows = d.ny * d.nzp;
    for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
        const int k = row / d.ny;
        const unsigned o = ((unsigned)row * (unsigned)d.nxpp + (unsigned)i) * (unsigned)sizeof(float);
        auto ld2 = [&](const float* p) {
            return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(p) + o);
        };
        double up[2][3], g[2][3][3];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            const float2 v = ld2(a.ui[l]);
            up[0][l] = (double)v.x;
            up[1][l] = (double)v.y;
        }
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const float2 v = ld2(a.g[l + 3 * m]);
                g[0][l][m] = (double)v.x;
                g[1][l][m] = (double)v.y;
            }
        float out[2][22];
        dynm_stage1_math(a, k, i, up[0], g[0], out[0]);
        dynm_stage1_math(a, k, i + 1, up[1], g[1], out[1]);
#pragma unroll
        for (int n = 0; n < 22; ++n)
            *reinterpret_cast<float2*>(reinterpret_cast<char*>(dynm_stage1_dst(a, n)) + o) =
                make_float2(out[0][n], out[1][n]);
    }
}

// stage 2 at one point of plane k from its 21 filtered values x: S' (6), L1 (6), u' (3),
// p (6); returns M:L and M:M
__device__ __forceinline__ void dynm_stage2_math(const DynMixArgs& a, int nzp, int k, const double (&x)[21],
                                                 float& num_o, float& den_o) {
    const double Ub = a.base[k];
    const double SUh = a.base[nzp + k];
    const double LU = a.base[2 * nzp + k];
    double sh[6], snh = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        sh[q] = x[q] + (q == 2 ? SUh : 0.0);
        const double w = (kL[q] == kM[q]) ? 1.0 : 2.0;
        snh = snh + w * 2.0 * sh[q] * sh[q];
    }
    snh = sqrt(snh);
    const double cd = 2.0 * a.delta[k];
    double uh[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) uh[l] = x[12 + l];
    const double Uh[3] = {Ub, 0.0, 0.0};
    double num = 0.0, den = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int l = kL[q], m = kM[q];
        const double mij = 2.0 * (x[6 + q] - cd * cd * snh * sh[q]);
        double lij = x[15 + q] - (Uh[l] * uh[m] + uh[l] * Uh[m] + uh[l] * uh[m]);
        if (q == 0) lij = lij + LU;
        const double w = (l == m) ? 1.0 : 2.0;
        num = num + w * mij * lij;
        den = den + w * mij * mij;
    }
    num_o = (float)num;
    den_o = (float)den;
}

__device__ __forceinline__ const float* dynm_stage2_src(const DynMixArgs& a, int n) {
    return n < 6 ? a.sph[n] : n < 12 ? a.l1h[n - 6] : n < 15 ? a.uh[n - 12] : a.ph[n - 15];
}

template <class Off>
__global__ void __launch_bounds__(256) k_dynm_stage2(DDims d, DynMixArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx) return;
    const int nrows = d.ny * d.nzp;
    for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
        const int k = row / d.ny;
        const Off o = ((Off)row * (Off)d.nxpp + (Off)i) * (Off)sizeof(float);
        double x[21];
#pragma unroll
        for (int n = 0; n < 21; ++n) x[n] = (double)ldf(dynm_stage2_src(a, n), o);
        float num, den;
        dynm_stage2_math(a, d.nzp, k, x, num, den);
        stf(a.num, o, num);
        stf(a.den, o, den);
    }
}

// The same with two adjacent x points per thread (8-byte loads and stores).
__global__ void __launch_bounds__(256) k_dynm_stage2x2(DDims d, DynMixArgs a) {
    const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.nx) return;
    const int nrows = d.ny * d.nzp;
    for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
        const int k = row / d.ny;
        const unsigned o = ((unsigned)row * (unsigned)d.nxpp + (unsigned)i) * (unsigned)sizeof(float);
        double x[2][21];
#pragma unroll
        for (int n = 0; n < 21; ++n) {
            const float2 v = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(dynm_stage2_src(a, n)) + o);
            x[0][n] = (double)v.x;
            x[1][n] = (double)v.y;
        }
        float num[2], den[2];
        dynm_stage2_math(a, d.nzp, k, x[0], num[0], den[0]);
        dynm_stage2_math(a, d.nzp, k, x[1], num[1], den[1]);
        *reinterpret_cast<float2*>(reinterpret_cast<char*>(a.num) + o) = make_float2(num[0], num[1]);
        *reinterpret_cast<float2*>(reinterpret_cast<char*>(a.den) + o) = make_float2(den[0], den[1]);
    }
}

// stage 3 at one point of plane k from the coefficient ratio ct = <M:L> / <M:M>
template <class Off>
__device__ __forceinline__ void dynm_stage3_point(const DynMixArgs& a, Off o, int k, double ct) {
    if (ct < 0.0 || isnan(ct)) ct = 0.0;
    if (ct > 0.2 * 0.2) ct = 0.2 * 0.2;
    const double dl = a.delta[k];
    const double nu = ct * dl * dl * (double)ldf(a.snorm, o);
    const double su = 0.5 * a.ubarp[k];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double s = (double)ldf(a.sp[q], o) + (q == 2 ? su : 0.0);
        const double tau = -2.0 * nu * s;
        stf(a.h[q], o, (float)(0.5 * (double)ldf(a.h[q], o) + tau));  // h holds p on entry
    }
}

template <class Off>
__global__ void __launch_bounds__(256) k_dynm_stage3(DDims d, DynMixArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx) return;
    const int nrows = d.ny * d.nzp;
    for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
        const int k = row / d.ny;
        const Off o = ((Off)row * (Off)d.nxpp + (Off)i) * (Off)sizeof(float);
        const double ct =
            a.cplane ? a.cplane[k] / a.cplane[d.nzp + k] : (double)ldf(a.numf, o) / (double)ldf(a.denf, o);
        dynm_stage3_point(a, o, k, ct);
    }
}

// The local (1, 1) average of M:L and M:M and stage 3 in one pass: k_filt3_lds's steps
// (a TX x R tile with periodic halos global -> registers -> LDS, the x pass into a second
// LDS array, the y pass, the z pass over three planes in registers, the next tile's loads
// in flight) for num and den in turn on every plane, and at each output point the two
// averages, rounded to FP32 as stored averages would be, go into stage 3.  Same
// expressions and order as the filter followed by k_dynm_stage3: bit-identical to them.
template <int R, int TX>
__global__ void __launch_bounds__(256) k_dynm_avg3(DDims d, DynMixArgs A, const double* __restrict__ w, int KC) {
    constexpr int NT = (R + 2) * (TX + 2);  // tile elements
    constexpr int NL = (NT + 255) / 256;    // tile loads per thread
    constexpr int NP = (R * TX) / 256;      // output points per thread
    static_assert((R * TX) % 256 == 0, "R * TX must be a multiple of 256");
    __shared__ float tile[NT];
    __shared__ double xs[(R + 2) * TX];
    const int ntx = (d.nx + TX - 1) / TX;
    const int i0 = (blockIdx.x % ntx) * TX;
    const int J0 = (blockIdx.x / ntx) * R;
    const int k0 = blockIdx.y * KC;
    const int k1 = min(k0 + KC, d.nzp);
    const float* src[2] = {A.num, A.den};
    const double a = 1.0, b = 1.0;
    const double s = 2. * a + b;
    const double s2 = 2.00 * a + b;
    const bool p2 = pow2_exact(s);
    const double is = 1.0 / s;
    const long plane = (long)d.ny * d.nxpp;
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
    float pre[NL];
    auto load = [&](int f, int k) {
        const float* qk = src[f] + (long)k * plane;
#pragma unroll
        for (int m = 0; m < NL; ++m) pre[m] = (toff[m] >= 0) ? qk[toff[m]] : 0.f;
    };
    // xy-filtered values of the tile in `pre`; then the loads of (fn, kn) (kn >= 0) in flight
    auto stage = [&](double (&y)[NP], int fn, int kn) {
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int e = threadIdx.x + 256 * m;
            if (e < NT) tile[e] = pre[m];
        }
        __syncthreads();
        if (kn >= 0) load(fn, kn);
        for (int e = threadIdx.x; e < (R + 2) * TX; e += 256) {
            const int rr = e / TX, c = e - rr * TX;
            const float* t = tile + rr * (TX + 2) + c;
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
    // planes k0 - 1 (if any) .. klast, num then den on each
    const int klast = min(k1, d.nzp - 1);
    const int kstart = (k0 > 0) ? k0 - 1 : k0;
    double Yp[2][NP], Yc[2][NP], Yn[2][NP];
    auto out = [&](int k, const double (&vn)[NP], const double (&vd)[NP]) {
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const int p = threadIdx.x + 256 * m;
            const int r = p / TX, c = p - r * TX;
            const int j = J0 + r, i = i0 + c;
            if (j < d.ny && i < d.nx) {
                const unsigned o = (unsigned)(((k * d.ny + j) * d.nxpp + i) * (int)sizeof(float));
                dynm_stage3_point(A, o, k, (double)(float)vn[m] / (double)(float)vd[m]);
            }
        }
    };
    load(0, kstart);
    for (int kk = kstart; kk <= klast; ++kk) {
        double y[2][NP];
        stage(y[0], 1, kk);
        stage(y[1], 0, (kk + 1 <= klast) ? kk + 1 : -1);
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
            for (int m = 0; m < NP; ++m) {
                if (kk < k0)
                    Yp[f][m] = y[f][m];
                else if (kk == k0)
                    Yc[f][m] = y[f][m];
                else
                    Yn[f][m] = y[f][m];
            }
        if (kk > k0) {
            const int k = kk - 1;
            const bool interior = (k > 0 && k < d.nzp - 1);
            const double w0 = w[k], w1 = w[d.nzp + k], w2 = w[2 * d.nzp + k];
            double v[2][NP];
#pragma unroll
            for (int f = 0; f < 2; ++f)
#pragma unroll
                for (int m = 0; m < NP; ++m) {
                    if (interior) {
                        double t = w0 * Yp[f][m] + w1 * Yc[f][m] + w2 * Yn[f][m];
                        v[f][m] = fdiv(t, s, is, p2);
                    } else {
                        v[f][m] = Yc[f][m];
                    }
                    Yp[f][m] = Yc[f][m];
                    Yc[f][m] = Yn[f][m];
                }
            out(k, v[0], v[1]);
        }
    }
    // the last plane of the grid keeps its xy-filtered value (no z pass)
    if (k1 == d.nzp) out(d.nzp - 1, Yc[0], Yc[1]);
}

int ph_dyn_mixed_avg3(lsb_ctx* c, const DynMixArgs& a, hipStream_t s) {
    const DDims d = ddims(c);
    // 32-bit element offsets; the (1, 1) weights on the device
    if ((double)d.ny * d.nzp * d.nxpp * sizeof(float) >= 2147483648.0 || a.cplane) return LSB_E_UNSUPPORTED;
    const double* w = c->d_fz + 3 * c->nzp;  // the (1, 1) interpolating weights
    constexpr int R = 4, TX = 128;
    const int ntx = (d.nx + TX - 1) / TX;
    const long tiles = (long)ntx * ((d.ny + R - 1) / R);
    // 16 planes per block, fewer (each block recomputes two halo planes) until the grid
    // holds some 2048 blocks (8 per CU)
    int KC = 16;
    while (KC > 4 && tiles * ((d.nzp + KC - 1) / KC) < 2048) KC /= 2;
    const dim3 grid((unsigned)tiles, (unsigned)((d.nzp + KC - 1) / KC));
    int tok = -1;
    // num, den read; snorm, sp (6), h (6) read; h (6) written
    prof_mark_begin(c, PC_NL, s, &tok, 0.0, 21.0 * 4.0 * (double)d.nx * d.ny * d.nzp);
    hipLaunchKernelGGL((k_dynm_avg3<R, TX>), grid, dim3(256), 0, s, d, a, w, KC);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// Plane means of two fields over the nx physical points of every row: B partial sums per
// plane (one block each), then the B partials of each plane summed in a fixed order.
constexpr int kPlaneBlocks = 16;

template <class T>
__global__ void __launch_bounds__(256) k_plane_partial2(DDims d, const T* __restrict__ f0, const T* __restrict__ f1,
                                                        double* __restrict__ part) {
    const int k = blockIdx.y;
    const int b = blockIdx.x;
    const long n = (long)d.nx * d.ny;
    double s0 = 0.0, s1 = 0.0;
    for (long t = (long)b * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % d.nx), j = (int)(t / d.nx);
        const long o = ((long)k * d.ny + j) * d.nxpp + i;
        s0 += (double)f0[o];
        s1 += (double)f1[o];
    }
    __shared__ double sh0[256], sh1[256];
    sh0[threadIdx.x] = s0;
    sh1[threadIdx.x] = s1;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sh0[threadIdx.x] += sh0[threadIdx.x + h];
            sh1[threadIdx.x] += sh1[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((long)k * kPlaneBlocks + b) * 2 + 0] = sh0[0];
        part[((long)k * kPlaneBlocks + b) * 2 + 1] = sh1[0];
    }
}

__global__ void k_plane_finish2(DDims d, const double* __restrict__ part, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.nzp) return;
    double s0 = 0.0, s1 = 0.0;
    for (int b = 0; b < kPlaneBlocks; ++b) {
        s0 += part[((long)k * kPlaneBlocks + b) * 2 + 0];
        s1 += part[((long)k * kPlaneBlocks + b) * 2 + 1];
    }
    const double n = (double)d.nx * (double)d.ny;
    out[k] = s0 / n;
    out[d.nzp + k] = s1 / n;
}

template <class T>
int ph_plane_mean2(lsb_ctx* c, const T* f0, const T* f1, double* scratch, double* out, hipStream_t s) {
    const DDims d = ddims(c);
    hipLaunchKernelGGL((k_plane_partial2<T>), dim3(kPlaneBlocks, (unsigned)d.nzp), dim3(256), 0, s, d, f0, f1,
                       scratch);
    LSB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_plane_finish2, dim3((unsigned)((d.nzp + 63) / 64)), dim3(64), 0, s, d, scratch, out);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
template int ph_plane_mean2<double>(lsb_ctx*, const double*, const double*, double*, double*, hipStream_t);
template int ph_plane_mean2<float>(lsb_ctx*, const float*, const float*, double*, double*, hipStream_t);

size_t ph_plane_mean2_scratch(const lsb_ctx* c) { return (size_t)c->nzp * kPlaneBlocks * 2; }

// LSB_NO_DYNX2=1: stages 1 and 2 one point per thread (for tests)
static bool stage1_single() {
    const char* e = std::getenv("LSB_NO_DYNX2");
    return e && e[0] == '1';
}

int ph_dyn_mixed(lsb_ctx* c, int stage, const DynMixArgs& a, hipStream_t s) {
    const DDims d = ddims(c);
    const dim3 g = row_grid(d);
    int tok = -1;
    // FP32 fields read + written per point by each stage
    const double per = (stage == 1) ? 12.0 + 22.0 : (stage == 2 ? 21.0 + 2.0 : 2.0 + 1.0 + 6.0 + 12.0);
    prof_mark_begin(c, PC_NL, s, &tok, 0.0, per * 4.0 * (double)d.nx * d.ny * d.nzp);
    // 32-bit byte offsets while a field is below 4 GiB
    const bool small = (double)d.ny * d.nzp * d.nxpp * sizeof(float) < 4294967296.0;
    switch (stage) {
        case 1:
            if (small && (d.nx % 2) == 0 && (d.nxpp % 2) == 0 && !stage1_single()) {
                const dim3 g2((unsigned)((d.nx / 2 + 255) / 256), g.y);
                hipLaunchKernelGGL(k_dynm_stage1x2, g2, dim3(256), 0, s, d, a);
            } else if (small)
                hipLaunchKernelGGL(k_dynm_stage1<unsigned>, g, dim3(256), 0, s, d, a);
            else
                hipLaunchKernelGGL(k_dynm_stage1<size_t>, g, dim3(256), 0, s, d, a);
            break;
        case 2:
            if (small && (d.nx % 2) == 0 && (d.nxpp % 2) == 0 && !stage1_single()) {
                const dim3 g2((unsigned)((d.nx / 2 + 255) / 256), g.y);
                hipLaunchKernelGGL(k_dynm_stage2x2, g2, dim3(256), 0, s, d, a);
            } else if (small)
                hipLaunchKernelGGL(k_dynm_stage2<unsigned>, g, dim3(256), 0, s, d, a);
            else
                hipLaunchKernelGGL(k_dynm_stage2<size_t>, g, dim3(256), 0, s, d, a);
            break;
        default:
            if (small)
                hipLaunchKernelGGL(k_dynm_stage3<unsigned>, g, dim3(256), 0, s, d, a);
            else
                hipLaunchKernelGGL(k_dynm_stage3<size_t>, g, dim3(256), 0, s, d, a);
            break;
    }
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// z-filtered base profiles of the mixed dynamic model (FP64, host): U^ = filter(ubar),
// (S^U)^_02 = filter(ubarp / 2), (U^2)^ - (U^)^2.  The x-y passes leave a function of z
// unchanged; the z pass uses the interpolating weights of interior planes (the boundary
// planes keep their value, filters.f90:450-453).
void dyn_base_profiles(const lsb_ctx* c, std::vector<double>& out) {
    const int n = c->nzp;
    const double s = 2.0 * c->cfg.a + c->cfg.b;
    const std::vector<double>& w = c->fz_ab;
    auto zf = [&](const std::vector<double>& v, int k) {
        if (k == 0 || k == n - 1) return v[k];
        return (w[k] * v[k - 1] + w[n + k] * v[k] + w[2 * n + k] * v[k + 1]) / s;
    };
    std::vector<double> u(n), su(n), u2(n);
    for (int k = 0; k < n; ++k) {
        u[k] = c->ubar[k];
        su[k] = 0.5 * c->ubarp[k];
        u2[k] = c->ubar[k] * c->ubar[k];
    }
    out.assign(3 * (size_t)n, 0.0);
    for (int k = 0; k < n; ++k) {
        const double ub = zf(u, k);
        out[k] = ub;
        out[n + k] = zf(su, k);
        out[2 * n + k] = zf(u2, k) - ub * ub;
    }
}

}  // namespace lsb
