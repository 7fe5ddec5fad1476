// Register-resident horizontal FFT passes for power-of-two lengths
// L = N1 * N2 in {16, 32, 64, 128, 256} (the general mixed-radix path in
// fft.hip covers every other length).
//
// Index split n = N2*n1 + n2, k = k1 + N1*k2:
//   stage 1  each thread (one n2) loads its N1 points straight from HBM and does
//            a length-N1 DFT in registers, then multiplies by W_L^(n2*k1);
//   exchange one write/read through a padded LDS tile (the N1 x N2 transpose);
//   stage 2  length-N2 DFTs in registers, results stored with unit-stride lanes.
// Loads and stores are coalesced (consecutive lanes touch consecutive complex
// words of one row, or of one row of a tile of kx columns).  The r2c split /
// c2r merge of the x pass use the natural-order spectrum staged once more in
// the same LDS tile.  Semantics are those of fft.hip (scaling, Nyquist drop,
// zero pads).
//
// Every pass is templated on the arithmetic type R and on the element types of
// its input (TI) and output (TO): the FP64 form uses double throughout, the
// mixed form reads FP64 spectra and writes FP32 physical fields (and back), the
// conversion happening in registers at the load/store.  The inverse y pass can
// apply d/dx or d/dy (x i*kx, x i*ky, derivatives.f90:261-303) to a field as it
// loads it, so spectral derivative fields never round-trip through HBM.

#include <cstring>

#include "lsb_internal.h"

namespace lsb {

// ------------------------------------------------------------- small DFTs
template <class R, int S>
__device__ __forceinline__ cpx<R> f_mul_si(cpx<R> a) {
    if (S < 0) return {a.im, -a.re};
    return {-a.im, a.re};
}

template <class R, int S>
__device__ __forceinline__ void f_dft4(cpx<R>& a0, cpx<R>& a1, cpx<R>& a2, cpx<R>& a3) {
    const cpx<R> s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = csub(a1, a3);
    const cpx<R> w = f_mul_si<R, S>(d13);
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, w);
    a3 = csub(d02, w);
}

// exp(S*2*pi*i*m/16) for m = 0..15
template <class R, int S>
__device__ __forceinline__ cpx<R> w16(int m) {
    const R c[16] = {(R)1.0L,
                     (R)0.92387953251128675612818318939678829L,
                     (R)0.70710678118654752440084436210484904L,
                     (R)0.38268343236508977172845998403039887L,
                     (R)0.0L,
                     (R)-0.38268343236508977172845998403039887L,
                     (R)-0.70710678118654752440084436210484904L,
                     (R)-0.92387953251128675612818318939678829L,
                     (R)-1.0L,
                     (R)-0.92387953251128675612818318939678829L,
                     (R)-0.70710678118654752440084436210484904L,
                     (R)-0.38268343236508977172845998403039887L,
                     (R)0.0L,
                     (R)0.38268343236508977172845998403039887L,
                     (R)0.70710678118654752440084436210484904L,
                     (R)0.92387953251128675612818318939678829L};
    const int ms = (m + 12) & 15;  // sin(x) = cos(x - pi/2)
    return {c[m & 15], (R)S * c[ms]};
}

template <class R, int S, int N>
__device__ __forceinline__ void f_dft(cpx<R>* x) {
    if constexpr (N == 2) {
        const cpx<R> a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (N == 4) {
        f_dft4<R, S>(x[0], x[1], x[2], x[3]);
    } else if constexpr (N == 8) {
        cpx<R> e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
        cpx<R> o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
        f_dft4<R, S>(e0, e1, e2, e3);
        f_dft4<R, S>(o0, o1, o2, o3);
        const cpx<R> w1 = cmul(o1, w16<R, S>(2));
        const cpx<R> w2 = cmul(o2, w16<R, S>(4));
        const cpx<R> w3 = cmul(o3, w16<R, S>(6));
        x[0] = cadd(e0, o0);
        x[4] = csub(e0, o0);
        x[1] = cadd(e1, w1);
        x[5] = csub(e1, w1);
        x[2] = cadd(e2, w2);
        x[6] = csub(e2, w2);
        x[3] = cadd(e3, w3);
        x[7] = csub(e3, w3);
    } else {  // 16 = 4 x 4
        cpx<R> t[16];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            cpx<R> a0 = x[b], a1 = x[4 + b], a2 = x[8 + b], a3 = x[12 + b];
            f_dft4<R, S>(a0, a1, a2, a3);
            t[4 * b + 0] = a0;
            t[4 * b + 1] = a1;
            t[4 * b + 2] = a2;
            t[4 * b + 3] = a3;
        }
#pragma unroll
        for (int b = 1; b < 4; ++b)
#pragma unroll
            for (int c = 1; c < 4; ++c) t[4 * b + c] = cmul(t[4 * b + c], w16<R, S>(b * c));
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            cpx<R> a0 = t[c], a1 = t[4 + c], a2 = t[8 + c], a3 = t[12 + c];
            f_dft4<R, S>(a0, a1, a2, a3);
            x[c] = a0;
            x[c + 4] = a1;
            x[c + 8] = a2;
            x[c + 12] = a3;
        }
    }
}

// stage 1 twiddle: v[k1] *= W_L^(n2 k1) with W_L = exp(S 2 pi i / L), table tw[m] = exp(-2 pi i m/L)
// (n2 = 0 multiplies by tw[0] = 1 exactly: no branch, so the N1 - 1 loads issue
// together instead of each waiting inside a divergent region)
template <class R, int S, int N1>
__device__ __forceinline__ void f_twiddle(cpx<R>* v, int n2, const cpx<R>* __restrict__ tw) {
    cpx<R> w[N1];
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) w[k1] = tw[n2 * k1];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        if (S > 0) w[k1].im = -w[k1].im;
        v[k1] = cmul(v[k1], w[k1]);
    }
}

// the same from a copy of the table in LDS (short latency: no batching needed)
template <class R, int S, int N1>
__device__ __forceinline__ void f_twiddle_lds(cpx<R>* v, int n2, const cpx<R>* tw) {
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        cpx<R> w = tw[n2 * k1];
        if (S > 0) w.im = -w.im;
        v[k1] = cmul(v[k1], w);
    }
}

// --------------------------------------------------------------- X passes
template <class R, class T>
__device__ __forceinline__ cpx<R> ld_c(const T* p, long i) {
    const T* q = p + 2 * i;
    return {(R)q[0], (R)q[1]};
}
template <class R, class T>
__device__ __forceinline__ void st_c(T* p, long i, cpx<R> v) {
    T* q = p + 2 * i;
    q[0] = (T)v.re;
    q[1] = (T)v.im;
}
template <>
__device__ __forceinline__ cpx<double> ld_c<double, double>(const double* p, long i) {
    return reinterpret_cast<const cpx<double>*>(p)[i];
}
template <>
__device__ __forceinline__ void st_c<double, double>(double* p, long i, cpx<double> v) {
    reinterpret_cast<cpx<double>*>(p)[i] = v;
}
template <>
__device__ __forceinline__ cpx<float> ld_c<float, float>(const float* p, long i) {
    return reinterpret_cast<const cpx<float>*>(p)[i];
}
template <>
__device__ __forceinline__ void st_c<float, float>(float* p, long i, cpx<float> v) {
    reinterpret_cast<cpx<float>*>(p)[i] = v;
}

// Forward r2c of rows: row r of nf*nrows, input stride istride reals (physical
// nxpp), output stride ostride reals (spectral nx).
template <class R, class TI, class TO, int N1, int N2>
__global__ void __launch_bounds__(128) k_xfwd_fast(FieldList<TI> in, FieldList<TO> out, int nrows, int nf,
                                                   int istride, int ostride, const cpx<R>* __restrict__ tw,
                                                   const cpx<R>* __restrict__ twr) {
    constexpr int L = N1 * N2;
    constexpr int LPB = 128 / N2;         // lines per block
    constexpr int TS = N1 * (N2 + 1);      // padded transpose tile per line (>= L)
    __shared__ cpx<R> T[LPB * TS];
    const int tid = threadIdx.x;
    const int l = tid / N2, n2 = tid % N2;
    const long r = (long)blockIdx.x * LPB + l;
    const long rlast = (long)nrows * nf - 1;
    const bool valid = r <= rlast;
    const long rc = valid ? r : rlast;  // lines past the end read the last line, store nothing
    const int f = (int)(rc / nrows);
    const long rr = rc - (long)f * nrows;
    const TI* src = in.p[f] + rr * istride;
    TO* dst = out.p[f] + rr * ostride;
    cpx<R> v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = ld_c<R>(src, N2 * n1 + n2);
    f_dft<R, -1, N1>(v);
    f_twiddle<R, -1, N1>(v, n2, tw);
    cpx<R>* Tl = T + l * TS;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) Tl[k1 * (N2 + 1) + n2] = v[k1];
    __syncthreads();
    cpx<R> u[N1 / N2][N2];
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q) {
        const int k1 = n2 + N2 * q;
#pragma unroll
        for (int j = 0; j < N2; ++j) u[q][j] = Tl[k1 * (N2 + 1) + j];
        f_dft<R, -1, N2>(u[q]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) Tl[(n2 + N2 * q) + N1 * k2] = u[q][k2];
    __syncthreads();
    if (!valid) return;
    // X_k = 1/2 (Z_k + conj Z_{L-k}) - i/2 W^k (Z_k - conj Z_{L-k})
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const int k = n2 + N2 * j;
        const cpx<R> zk = Tl[k];
        const cpx<R> zc = cconj(Tl[(L - k) & (L - 1)]);
        const cpx<R> a = {(R)0.5 * (zk.re + zc.re), (R)0.5 * (zk.im + zc.im)};
        const cpx<R> d = {(R)0.5 * (zk.re - zc.re), (R)0.5 * (zk.im - zc.im)};
        const cpx<R> wd = cmul(twr[k], d);
        st_c<R>(dst, k, cpx<R>{a.re + wd.im, a.im - wd.re});
    }
}

// Inverse c2r of rows holding L complex coefficients (input row stride istride
// reals) into physical rows (stride nxpp, the two pad reals written as zero);
// in place when in == out and istride == nxpp.  STAGE: the block's LPB rows, pads
// included, are one contiguous run of LPB*nxpp reals of one field (nrows a multiple of
// LPB); they are assembled in the LDS tile and stored as 16-byte vectors, so the
// stores fill whole 128-byte segments although the rows of stride nxpp = L*2 + 2 start
// off the segment grid.
template <class R, class TI, class TO, int N1, int N2, bool STAGE = false>
__global__ void __launch_bounds__(128) k_xinv_fast(FieldList<TI> in, FieldList<TO> out, int nrows, int nf,
                                                   int istride, int nxpp, const cpx<R>* __restrict__ tw,
                                                   const cpx<R>* __restrict__ twr) {
    constexpr int L = N1 * N2;
    constexpr int LPB = 128 / N2;
    constexpr int TS = N1 * (N2 + 1);
    __shared__ __align__(16) cpx<R> T[LPB * TS];  // 16-byte aligned for the staged vector copy
    const int tid = threadIdx.x;
    const int l = tid / N2, n2 = tid % N2;
    const long r = (long)blockIdx.x * LPB + l;
    const long rlast = (long)nrows * nf - 1;
    const bool valid = r <= rlast;
    // lines past the end read the last line (unconditional loads issue back to
    // back; a load under a per-element branch feeding the LDS store would wait
    // for each in turn) and store nothing
    const long rc = valid ? r : rlast;
    const int f = (int)(rc / nrows);
    const long rr = rc - (long)f * nrows;
    const TI* rin = in.p[f] + rr * istride;
    TO* row = out.p[f] + rr * nxpp;
    cpx<R>* Tl = T + l * TS;
    {
        cpx<R> t[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) t[j] = ld_c<R>(rin, n2 + N2 * j);
#pragma unroll
        for (int j = 0; j < N1; ++j) Tl[n2 + N2 * j] = t[j];
    }
    __syncthreads();
    // Z'_n = (X_n + conj X_{L-n}) + i W^-n (X_n - conj X_{L-n}); X_L = 0, Im X_0 ignored
    cpx<R> v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int n = N2 * n1 + n2;
        cpx<R> xk = Tl[n];
        cpx<R> xc;
        if (n == 0) {
            xk.im = 0;
            xc = {0, 0};
        } else {
            xc = cconj(Tl[L - n]);
        }
        const cpx<R> s = cadd(xk, xc);
        const cpx<R> d = csub(xk, xc);
        const cpx<R> wd = cmul(cconj(twr[n]), d);
        v[n1] = {s.re - wd.im, s.im + wd.re};
    }
    f_dft<R, 1, N1>(v);
    f_twiddle<R, 1, N1>(v, n2, tw);
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) Tl[k1 * (N2 + 1) + n2] = v[k1];
    __syncthreads();
    if constexpr (STAGE && sizeof(TO) <= sizeof(R)) {
        static_assert(LPB * TS * sizeof(cpx<R>) >= LPB * (2 * L + 2) * sizeof(TO), "staging tile");
        cpx<R> u[N1 / N2][N2];
#pragma unroll
        for (int q = 0; q < N1 / N2; ++q) {
            const int k1 = n2 + N2 * q;
#pragma unroll
            for (int j = 0; j < N2; ++j) u[q][j] = Tl[k1 * (N2 + 1) + j];
            f_dft<R, 1, N2>(u[q]);
        }
        __syncthreads();
        // the block's rows in their memory layout (stride nxpp = 2L + 2 reals of TO)
        TO* S = reinterpret_cast<TO*>(T);
        TO* Sl = S + (long)l * (2 * L + 2);
#pragma unroll
        for (int q = 0; q < N1 / N2; ++q) {
            const int k1 = n2 + N2 * q;
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                Sl[2 * (k1 + N1 * k2)] = (TO)u[q][k2].re;
                Sl[2 * (k1 + N1 * k2) + 1] = (TO)u[q][k2].im;
            }
        }
        if (n2 == 0) {
            Sl[2 * L] = (TO)0;
            Sl[2 * L + 1] = (TO)0;
        }
        __syncthreads();
        // the valid rows of the block: one run of the field (the host guarantees it)
        const long r0 = (long)blockIdx.x * LPB;
        const long nv = min((long)LPB, rlast + 1 - r0);
        const int f0 = (int)(r0 / nrows);
        TO* g = out.p[f0] + (r0 - (long)f0 * nrows) * nxpp;
        const long nreal = nv * nxpp;
        constexpr int VW = 16 / (int)sizeof(TO);
        using V = typename std::conditional<sizeof(TO) == 8, double2, float4>::type;
        const long nvec = nreal / VW;
        for (long e = tid; e < nvec; e += 128)
            reinterpret_cast<V*>(g)[e] = reinterpret_cast<const V*>(S)[e];
        for (long e = nvec * VW + tid; e < nreal; e += 128) g[e] = S[e];
        return;
    }
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q) {
        const int k1 = n2 + N2 * q;
        cpx<R> u[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) u[j] = Tl[k1 * (N2 + 1) + j];
        f_dft<R, 1, N2>(u);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) st_c<R>(row, k1 + N1 * k2, u[k2]);
    }
    if (n2 == 0) st_c<R>(row, L, cpx<R>{0, 0});  // the two pad reals
}

// --------------------------------------------------------------- Y passes
// Tiles of C kx-columns x L ky-rows of one plane; lane index c fastest.
// Forward: input rows of lh complex (stride lh), output the same layout, scaled;
// in place when in == out.
template <class R, class TI, class TO, int N1, int N2, int C>
__global__ void __launch_bounds__(C * N2) k_yfwd_fast(FieldList<TI> in, FieldList<TO> out, int nplanes, int lh,
                                                      const cpx<R>* __restrict__ tw, R scale) {
    constexpr int TS = N1 * (N2 + 1) + 1;
    __shared__ cpx<R> T[C * TS];
    const int tid = threadIdx.x;
    const int c = tid % C, n2 = tid / C;
    const int ntile = (lh + C - 1) / C;
    const int tile = blockIdx.x % ntile;
    const long plane = blockIdx.x / ntile;
    const int f = (int)(plane / nplanes);
    const long pk = plane - (long)f * nplanes;
    const TI* src = in.p[f] + pk * (long)(N1 * N2) * (2 * lh);
    TO* dst = out.p[f] + pk * (long)(N1 * N2) * (2 * lh);
    const int m = tile * C + c;
    const bool valid = m < lh;
    const int mc = valid ? m : lh - 1;  // columns past the end read the last one, store nothing
    cpx<R> v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = ld_c<R>(src, (long)(N2 * n1 + n2) * lh + mc);
    f_dft<R, -1, N1>(v);
    f_twiddle<R, -1, N1>(v, n2, tw);
    cpx<R>* Tc = T + c * TS;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) Tc[k1 * (N2 + 1) + n2] = v[k1];
    __syncthreads();
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q) {
        const int k1 = n2 + N2 * q;
        cpx<R> u[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) u[j] = Tc[k1 * (N2 + 1) + j];
        f_dft<R, -1, N2>(u);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            st_c<R>(dst, (long)(k1 + N1 * k2) * lh + m, cpx<R>{u[k2].re * scale, u[k2].im * scale});
    }
}

// Inverse: spectral rows (stride lh) in, rows of stride ph complex out.  ops.op[f]
// selects d/dx (x i*wavx) or d/dy (x i*wavy) applied to field f as it is loaded,
// in double before the conversion to R: (re, im) -> (-k im, k re).
template <class R, class TI, class TO, int N1, int N2, int C>
__global__ void __launch_bounds__(C * N2) k_yinv_fast(FieldList<TI> in, FieldList<TO> out, int nplanes, int lh,
                                                      int ph, const cpx<R>* __restrict__ tw, FftOps ops,
                                                      const double* __restrict__ wavx,
                                                      const double* __restrict__ wavy) {
    constexpr int TS = N1 * (N2 + 1) + 1;
    __shared__ cpx<R> T[C * TS];
    const int tid = threadIdx.x;
    const int c = tid % C, n2 = tid / C;
    const int ntile = (lh + C - 1) / C;
    const int tile = blockIdx.x % ntile;
    const long plane = blockIdx.x / ntile;
    const int f = (int)(plane / nplanes);
    const long pk = plane - (long)f * nplanes;
    const TI* src = in.p[f] + pk * (long)(N1 * N2) * (2 * lh);
    TO* dst = out.p[f] + pk * (long)(N1 * N2) * (2 * ph);
    const int m = tile * C + c;
    const bool valid = m < lh;
    const int mc = valid ? m : lh - 1;  // columns past the end read the last one, store nothing
    const int op = ops.op[f];
    TI qre[N1], qim[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const TI* q = src + 2 * ((long)(N2 * n1 + n2) * lh + mc);
        qre[n1] = q[0];
        qim[n1] = q[1];
    }
    cpx<R> v[N1];
    if (op == FOP_NONE) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = {(R)qre[n1], (R)qim[n1]};
    } else {
        const double kx = (op == FOP_DX) ? wavx[2 * mc] : 0.0;
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const double k = (op == FOP_DX) ? kx : wavy[N2 * n1 + n2];
            const double re = (double)qre[n1], im = (double)qim[n1];
            v[n1] = {(R)(-k * im), (R)(k * re)};
        }
    }
    f_dft<R, 1, N1>(v);
    f_twiddle<R, 1, N1>(v, n2, tw);
    cpx<R>* Tc = T + c * TS;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) Tc[k1 * (N2 + 1) + n2] = v[k1];
    __syncthreads();
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q) {
        const int k1 = n2 + N2 * q;
        cpx<R> u[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) u[j] = Tc[k1 * (N2 + 1) + j];
        f_dft<R, 1, N2>(u);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) st_c<R>(dst, (long)(k1 + N1 * k2) * ph + m, u[k2]);
    }
}

// Inverse y pass of a velocity spectrum and its x and y derivatives from one load:
// out.p[3 f] = F^-1_y(u), out.p[3 f + 1] = i kx F^-1_y(u) (kx is constant along a
// y column, so the x derivative is applied at the store), out.p[3 f + 2] =
// F^-1_y(i ky u) (ky is applied at the load).  Outputs are rows of stride ph.
template <class R, class TI, class TO, int N1, int N2, int C>
__global__ void __launch_bounds__(C * N2) k_yinv3_fast(FieldList<TI> in, FieldList<TO> out, int nplanes, int lh,
                                                       int ph, const cpx<R>* __restrict__ tw,
                                                       const double* __restrict__ wavx,
                                                       const double* __restrict__ wavy,
                                                       const double* __restrict__ ftab) {
    constexpr int TS = N1 * (N2 + 1) + 1;
    __shared__ cpx<R> T[C * TS];
    const int tid = threadIdx.x;
    const int c = tid % C, n2 = tid / C;
    const int ntile = (lh + C - 1) / C;
    const int tile = blockIdx.x % ntile;
    const long plane = blockIdx.x / ntile;
    const int f = (int)(plane / nplanes);
    const long pk = plane - (long)f * nplanes;
    const TI* src = in.p[f] + pk * (long)(N1 * N2) * (2 * lh);
    const long ooff = pk * (long)(N1 * N2) * (2 * ph);
    TO* du = out.p[3 * f] + ooff;
    TO* dx = out.p[3 * f + 1] + ooff;
    TO* dy = out.p[3 * f + 2] + ooff;
    const int m = tile * C + c;
    const bool valid = m < lh;
    const int mc = valid ? m : lh - 1;  // columns past the end read the last one, store nothing
    const double kx = wavx[2 * mc];
    double re[N1], im[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const TI* q = src + 2 * ((long)(N2 * n1 + n2) * lh + mc);
        re[n1] = (double)q[0];
        im[n1] = (double)q[1];
    }
    if (ftab) {
        // the spectral filter as a real factor per mode: the reference's
        // magnitude/phase rebuild (specfilter_mode) equals it up to the rounding
        // of re/|u|, i.e. ~eps |re|^2/|im| on modes with |im| << |re|
        double fac[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) fac[n1] = ftab[(N2 * n1 + n2) * lh + mc];
        __builtin_amdgcn_sched_barrier(0);  // keep the loads together, ahead of their uses
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            re[n1] *= fac[n1];
            im[n1] *= fac[n1];
        }
    }
    cpx<R>* Tc = T + c * TS;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        cpx<R> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            if (pass == 0) {
                v[n1] = {(R)re[n1], (R)im[n1]};
            } else {
                const double k = wavy[N2 * n1 + n2];
                v[n1] = {(R)(-k * im[n1]), (R)(k * re[n1])};
            }
        }
        f_dft<R, 1, N1>(v);
        f_twiddle<R, 1, N1>(v, n2, tw);
        if (pass == 1) __syncthreads();  // the first pass has read its tile
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) Tc[k1 * (N2 + 1) + n2] = v[k1];
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int q = 0; q < N1 / N2; ++q) {
                const int k1 = n2 + N2 * q;
                cpx<R> u[N2];
#pragma unroll
                for (int j = 0; j < N2; ++j) u[j] = Tc[k1 * (N2 + 1) + j];
                f_dft<R, 1, N2>(u);
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    const long o = (long)(k1 + N1 * k2) * ph + m;
                    if (pass == 0) {
                        st_c<R>(du, o, u[k2]);
                        st_c<R>(dx, o, cpx<R>{(R)(-kx) * u[k2].im, (R)kx * u[k2].re});
                    } else {
                        st_c<R>(dy, o, u[k2]);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------- dispatch
bool fast_fft_len(int L) { return L == 16 || L == 32 || L == 64 || L == 128 || L == 256; }

#define LSB_FAST_CASES(X) \
    X(16, 4, 4)           \
    X(32, 8, 4)           \
    X(64, 8, 8)           \
    X(128, 16, 8)         \
    X(256, 16, 16)

template <class R, class TI, class TO>
int fast_xfwd(const FieldList<TI>& in, const FieldList<TO>& out, int nrows, int nf, int lh, int istride,
              int ostride, const cpx<R>* tw, const cpx<R>* twr, hipStream_t s) {
    const long rows = (long)nrows * nf;
    if (rows <= 0) return LSB_OK;
    switch (lh) {
#define X(LL, A, B)                                                                                                 \
    case LL:                                                                                                        \
        hipLaunchKernelGGL((k_xfwd_fast<R, TI, TO, A, B>), dim3((unsigned)((rows + 128 / B - 1) / (128 / B))),       \
                           dim3(128), 0, s, in, out, nrows, nf, istride, ostride, tw, twr);                         \
        break;
        LSB_FAST_CASES(X)
#undef X
        default: return LSB_E_ARG;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

template <class R, class TI, class TO>
int fast_xinv(const FieldList<TI>& in, const FieldList<TO>& out, int nrows, int nf, int lh, int istride, int nxpp,
              const cpx<R>* tw, const cpx<R>* twr, hipStream_t s) {
    const long rows = (long)nrows * nf;
    if (rows <= 0) return LSB_OK;
    // the staged store: whole blocks of rows inside one field, rows of exactly 2 lh + 2
    // reals, fields 16-byte aligned
    bool aligned = nxpp == 2 * lh + 2 && sizeof(TO) <= sizeof(R);
    const char* off = std::getenv("LSB_NO_XSTAGE");  // the direct row stores, for tests
    if (off && off[0] && std::strcmp(off, "0") != 0) aligned = false;
    for (int f = 0; f < nf && aligned; ++f) aligned = ((uintptr_t)out.p[f] & 15) == 0;
    switch (lh) {
#define X(LL, A, B)                                                                                                 \
    case LL:                                                                                                        \
        if (aligned && nrows % (128 / B) == 0)                                                                      \
            hipLaunchKernelGGL((k_xinv_fast<R, TI, TO, A, B, true>),                                                \
                               dim3((unsigned)((rows + 128 / B - 1) / (128 / B))), dim3(128), 0, s, in, out, nrows, \
                               nf, istride, nxpp, tw, twr);                                                         \
        else                                                                                                        \
            hipLaunchKernelGGL((k_xinv_fast<R, TI, TO, A, B>), dim3((unsigned)((rows + 128 / B - 1) / (128 / B))),   \
                               dim3(128), 0, s, in, out, nrows, nf, istride, nxpp, tw, twr);                        \
        break;
        LSB_FAST_CASES(X)
#undef X
        default: return LSB_E_ARG;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// kx columns per y-pass tile: 128-byte row segments of the arithmetic type
template <class R>
constexpr int ytile_cols() {
    return 128 / (int)sizeof(cpx<R>);
}

template <class R, class TI, class TO, int C>
int fast_yfwd_c(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny,
                const cpx<R>* tw, R scale, hipStream_t s);

// kx columns per tile of a y pass: 8, 16 or (FP32 arithmetic) 32, from the context's
// tuning table; anything else gives the built-in default.  Each column is transformed
// by the same lanes in the same order whatever the tile width, so every choice gives
// the same bits.
template <class R>
static int ycols(int cols, int dflt) {
    if (cols == 8 || cols == 16 || (cols == 32 && sizeof(R) == 4)) return cols;
    return dflt;
}
#define LSB_YCOLS_DISPATCH(R, CALL, cols, dflt)     \
    switch (ycols<R>(cols, dflt)) {                 \
        case 8: return CALL(8);                     \
        case 32:                                    \
            if constexpr (sizeof(R) == 4) return CALL(32); \
            return LSB_E_ARG;                       \
        default: return CALL(16);                   \
    }

// The forward y pass defaults to 16 kx-columns per tile in both precisions: 256-byte row
// segments in FP64 (measured 0.57 -> 0.62 of the HBM rate against 8-column tiles on the
// 512 x 193 x 256 step, results unchanged), 128-byte ones in FP32.
template <class R, class TI, class TO>
int fast_yfwd(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny,
              const cpx<R>* tw, R scale, hipStream_t s, int cols) {
#define CALL(C) fast_yfwd_c<R, TI, TO, C>(in, out, nplanes, nf, lh, ny, tw, scale, s)
    LSB_YCOLS_DISPATCH(R, CALL, cols, 16)
#undef CALL
}

template <class R, class TI, class TO, int C>
int fast_yfwd_c(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny,
                const cpx<R>* tw, R scale, hipStream_t s) {
    const long nblk = (long)((lh + C - 1) / C) * nplanes * nf;
    if (nblk <= 0) return LSB_OK;
    switch (ny) {
#define X(LL, A, B)                                                                                          \
    case LL:                                                                                                 \
        hipLaunchKernelGGL((k_yfwd_fast<R, TI, TO, A, B, C>), dim3((unsigned)nblk), dim3(C * B), 0, s, in, out, \
                           nplanes, lh, tw, scale);                                                          \
        break;
        LSB_FAST_CASES(X)
#undef X
        default: return LSB_E_ARG;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

template <class R, class TI, class TO, int C>
static int fast_yinv_c(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny,
                       int ph, const cpx<R>* tw, const FftOps& ops, const double* wavx, const double* wavy,
                       hipStream_t s) {
    const long nblk = (long)((lh + C - 1) / C) * nplanes * nf;
    if (nblk <= 0) return LSB_OK;
    switch (ny) {
#define X(LL, A, B)                                                                                          \
    case LL:                                                                                                 \
        hipLaunchKernelGGL((k_yinv_fast<R, TI, TO, A, B, C>), dim3((unsigned)nblk), dim3(C * B), 0, s, in, out, \
                           nplanes, lh, ph, tw, ops, wavx, wavy);                                            \
        break;
        LSB_FAST_CASES(X)
#undef X
        default: return LSB_E_ARG;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

template <class R, class TI, class TO>
int fast_yinv(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nf, int lh, int ny, int ph,
              const cpx<R>* tw, const FftOps& ops, const double* wavx, const double* wavy, hipStream_t s,
              int cols) {
#define CALL(C) fast_yinv_c<R, TI, TO, C>(in, out, nplanes, nf, lh, ny, ph, tw, ops, wavx, wavy, s)
    LSB_YCOLS_DISPATCH(R, CALL, cols, ytile_cols<R>())
#undef CALL
}

template <class R, class TI, class TO, int C>
static int fast_yinv3_c(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nsrc, int lh, int ny,
                        int ph, const cpx<R>* tw, const double* wavx, const double* wavy, const double* ftab,
                        hipStream_t s) {
    const long nblk = (long)((lh + C - 1) / C) * nplanes * nsrc;
    if (nblk <= 0) return LSB_OK;
    switch (ny) {
#define X(LL, A, B)                                                                                           \
    case LL:                                                                                                  \
        hipLaunchKernelGGL((k_yinv3_fast<R, TI, TO, A, B, C>), dim3((unsigned)nblk), dim3(C * B), 0, s, in, out, \
                           nplanes, lh, ph, tw, wavx, wavy, ftab);                                            \
        break;
        LSB_FAST_CASES(X)
#undef X
        default: return LSB_E_ARG;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

template <class R, class TI, class TO>
int fast_yinv3(const FieldList<TI>& in, const FieldList<TO>& out, int nplanes, int nsrc, int lh, int ny, int ph,
               const cpx<R>* tw, const double* wavx, const double* wavy, const double* ftab, hipStream_t s,
               int cols) {
#define CALL(C) fast_yinv3_c<R, TI, TO, C>(in, out, nplanes, nsrc, lh, ny, ph, tw, wavx, wavy, ftab, s)
    LSB_YCOLS_DISPATCH(R, CALL, cols, ytile_cols<R>())
#undef CALL
}

// Forward y pass fused with the velocity right-hand side (one block: C kx-columns
// of one plane, one component l).  The three pieces are transformed in turn
// through the same LDS tile; N accumulates in FP64 in the order of the separate
// passes (nlt, then - i kx h_x, then - i ky h_y, each piece rounded to R after
// scaling as the separate forward pass stores it), so results are bitwise those of
// fast_yfwd followed by the right-hand-side kernel.
template <class TN>
__device__ __forceinline__ void ld_h(const void* p, long q, double& re, double& im) {
    const TN* t = reinterpret_cast<const TN*>(p) + 2 * q;
    re = (double)t[0];
    im = (double)t[1];
}
template <class TN>
__device__ __forceinline__ void st_h(void* p, long q, double re, double im) {
    TN* t = reinterpret_cast<TN*>(p) + 2 * q;
    t[0] = (TN)re;
    t[1] = (TN)im;
}

template <class R, class TH, int N1, int N2, int C>
__global__ void __launch_bounds__(C * N2) k_yfwd_rhs(FieldList<R> pc, int nplanes, int lh,
                                                     const cpx<R>* __restrict__ tw, R scale, YRhsArgs a) {
    constexpr int L = N1 * N2;
    constexpr int TS = N1 * (N2 + 1) + 1;
    constexpr int NQ = N1 / N2;
    __shared__ cpx<R> T[C * TS];
    __shared__ cpx<R> Tw[L];
    const int tid = threadIdx.x;
    for (int t = tid; t < L; t += C * N2) Tw[t] = tw[t];
    const int c = tid % C, n2 = tid / C;
    const int ntile = (lh + C - 1) / C;
    const int tile = blockIdx.x % ntile;
    const long rest = blockIdx.x / ntile;
    const int pk = (int)(rest % nplanes);
    const int l = (int)(rest / nplanes);
    const int m = tile * C + c;
    const bool valid = m < lh;
    const int mc = valid ? m : lh - 1;  // columns past the end read the last one, store nothing
    const double wx = a.wavx[2 * mc];
    double nr[NQ][N2], ni[NQ][N2];
    cpx<R>* Tc = T + c * TS;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        // one piece at a time: the next piece's loads are not hoisted over this
        // piece's second stage (three live tiles would halve the occupancy)
        __builtin_amdgcn_sched_barrier(0);
        const R* src = pc.p[3 * l + q] + (long)pk * L * (2 * lh);
        cpx<R> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = ld_c<R>(src, (long)(N2 * n1 + n2) * lh + mc);
        f_dft<R, -1, N1>(v);
        __syncthreads();  // the twiddle table is in place; the previous piece has read its tile
        f_twiddle_lds<R, -1, N1>(v, n2, Tw);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) Tc[k1 * (N2 + 1) + n2] = v[k1];
        __syncthreads();
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) {
            const int k1 = n2 + N2 * qq;
            cpx<R> u[N2];
#pragma unroll
            for (int j = 0; j < N2; ++j) u[j] = Tc[k1 * (N2 + 1) + j];
            f_dft<R, -1, N2>(u);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const double vr = (double)(R)(u[k2].re * scale), vi = (double)(R)(u[k2].im * scale);
                if (q == 0) {
                    nr[qq][k2] = vr;
                    ni[qq][k2] = vi;
                } else if (q == 1) {
                    nr[qq][k2] = nr[qq][k2] - (-wx * vi);
                    ni[qq][k2] = ni[qq][k2] - (wx * vr);
                } else {
                    const double wy = a.wavy[k1 + N1 * k2];
                    nr[qq][k2] = nr[qq][k2] - (-wy * vi);
                    ni[qq][k2] = ni[qq][k2] - (wy * vr);
                }
            }
        }
    }
    if (!valid) return;
    const double ub = (l == 0 && m == 0) ? a.ubarpp[pk] : 0.0;
#pragma unroll
    for (int qq = 0; qq < NQ; ++qq) {
        const int k1 = n2 + N2 * qq;
#pragma unroll
        for (int b0 = 0; b0 < N2; b0 += 4) {
            // the history of four outputs is loaded before their stores (and not
            // earlier: the barrier bounds the live registers)
            __builtin_amdgcn_sched_barrier(0);
            double h1r[4], h1i[4], h2r[4], h2i[4], u0r[4], u0i[4], u1r[4], u1i[4], u2r[4], u2i[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int ky = k1 + N1 * (b0 + b);
                const long qd = ((long)pk * L + ky) * lh + m;
                ld_h<TH>(a.nl[l], qd, h1r[b], h1i[b]);
                ld_h<TH>(a.nlm[l], qd, h2r[b], h2i[b]);
                ld_h<double>(a.u[l], qd, u0r[b], u0i[b]);
                ld_h<double>(a.un[l], qd, u1r[b], u1i[b]);
                ld_h<double>(a.unm[l], qd, u2r[b], u2i[b]);
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int k2 = b0 + b;
                const int ky = k1 + N1 * k2;
                const long qd = ((long)pk * L + ky) * lh + m;
                double xr = nr[qq][k2], xi = ni[qq][k2];
                if (sizeof(TH) < sizeof(double)) {
                    xr = (double)(TH)xr;
                    xi = (double)(TH)xi;
                }
                st_h<TH>(a.nout[l], qd, xr, xi);
                const double cr = a.ab0 * xr + a.ab1 * h1r[b] + a.ab2 * h2r[b];
                const double ci = a.ab0 * xi + a.ab1 * h1i[b] + a.ab2 * h2i[b];
                double rr = -a.dtnui * (a.dt * cr + a.bdf0 * u0r[b] + a.bdf1 * u1r[b] + a.bdf2 * u2r[b]);
                const double ri = -a.dtnui * (a.dt * ci + a.bdf0 * u0i[b] + a.bdf1 * u1i[b] + a.bdf2 * u2i[b]);
                if (ky == 0) rr = rr - ub;
                st_h<double>(a.rhs[l], qd, rr, ri);
            }
        }
    }
}

template <class R, class TH, int C>
static int fast_yfwd_rhs_c(const FieldList<R>& pieces, int nplanes, int lh, int ny, const cpx<R>* tw, R scale,
                           const YRhsArgs& a, hipStream_t s) {
    const long nblk = (long)((lh + C - 1) / C) * nplanes * 3;
    if (nblk <= 0) return LSB_OK;
    switch (ny) {
#define X(LL, A, B)                                                                                            \
    case LL:                                                                                                   \
        hipLaunchKernelGGL((k_yfwd_rhs<R, TH, A, B, C>), dim3((unsigned)nblk), dim3(C * B), 0, s, pieces, nplanes, \
                           lh, tw, scale, a);                                                                  \
        break;
        LSB_FAST_CASES(X)
#undef X
        default: return LSB_E_ARG;
    }
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
template <class R, class TH>
int fast_yfwd_rhs(const FieldList<R>& pieces, int nplanes, int lh, int ny, const cpx<R>* tw, R scale,
                  const YRhsArgs& a, hipStream_t s, int cols) {
#define CALL(C) fast_yfwd_rhs_c<R, TH, C>(pieces, nplanes, lh, ny, tw, scale, a, s)
    LSB_YCOLS_DISPATCH(R, CALL, cols, ytile_cols<R>())
#undef CALL
}
template int fast_yfwd_rhs<double, double>(const FieldList<double>&, int, int, int, const cpx<double>*, double,
                                           const YRhsArgs&, hipStream_t, int);
template int fast_yfwd_rhs<float, float>(const FieldList<float>&, int, int, int, const cpx<float>*, float,
                                         const YRhsArgs&, hipStream_t, int);

#define LSB_INST_FAST(R, TI, TO)                                                                                  \
    template int fast_xfwd<R, TI, TO>(const FieldList<TI>&, const FieldList<TO>&, int, int, int, int, int,        \
                                      const cpx<R>*, const cpx<R>*, hipStream_t);                                 \
    template int fast_xinv<R, TI, TO>(const FieldList<TI>&, const FieldList<TO>&, int, int, int, int, int,        \
                                      const cpx<R>*, const cpx<R>*, hipStream_t);                                 \
    template int fast_yfwd<R, TI, TO>(const FieldList<TI>&, const FieldList<TO>&, int, int, int, int,             \
                                      const cpx<R>*, R, hipStream_t, int);                                        \
    template int fast_yinv<R, TI, TO>(const FieldList<TI>&, const FieldList<TO>&, int, int, int, int, int,        \
                                      const cpx<R>*, const FftOps&, const double*, const double*, hipStream_t,  \
                                      int);                                                                       \
    template int fast_yinv3<R, TI, TO>(const FieldList<TI>&, const FieldList<TO>&, int, int, int, int, int,       \
                                       const cpx<R>*, const double*, const double*, const double*, hipStream_t,  \
                                       int);

LSB_INST_FAST(double, double, double)
LSB_INST_FAST(float, float, float)    // mixed: x passes (physical <-> compact staging)
LSB_INST_FAST(float, double, float)   // mixed: inverse y pass (FP64 spectrum -> FP32 staging)
LSB_INST_FAST(float, float, double)   // mixed: forward y pass (FP32 staging -> FP64 spectrum)

}  // namespace lsb
