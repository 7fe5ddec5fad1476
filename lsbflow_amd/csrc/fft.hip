// Horizontal 2-D FFTs (replaces horfft.f90:61-150 and the FFTW plans of
// modfftw.f90:59-149).
//
// Layout (one rank): physical rows of nxpp = nx+2 reals (x fastest), spectral
// rows of nx reals = nx/2 complex (kx = 0..nx/2-1; the kx-Nyquist mode is not
// stored, horfft.f90:303-315), ny rows per z-plane, nzp planes per field.
//
// Forward (is = -1): X pass  physical row -> spectral row (r2c as a length
//   nx/2 complex FFT + split), then Y pass in place over tiles of kx columns
//   (c2c along ky, stride nx/2 complex) with the 1/(nx*ny) scale fused.
// Inverse (is = +1): Y pass spectral -> physical buffer (complex rows of
//   stride nxpp/2), then X pass in place (c2r; Im of the kx = 0 term and the
//   missing Nyquist term are ignored, as FFTW's c2r does).  Pads are zeroed.
//
// Each workgroup keeps its lines in LDS and runs a mixed-radix Stockham
// sequence; twiddles come from tables computed in long double on the host.

#include <cmath>

#include <cstdio>
#include <cstdlib>

#include "lsb_internal.h"

namespace lsb {

// ---------------------------------------------------------------- plans
bool make_plan(int L, FftPlan1D* plan) {
    plan->L = L;
    plan->nstage = 0;
    int n = L;
    // radix-8 first, then 4, then 2, 3, 5
    const int order[5] = {8, 4, 2, 3, 5};
    for (int oi = 0; oi < 5 && n > 1; ++oi) {
        int r = order[oi];
        while (n % r == 0 && n > 1) {
            if (plan->nstage >= kMaxStages) return false;
            plan->radix[plan->nstage++] = r;
            n /= r;
        }
    }
    return n == 1 && L >= 1;
}

template <class R>
static int make_twiddles(void** dptr, int L, int denom, int count) {
    // count entries exp(-2 pi i m / denom), m = 0..count-1
    std::vector<R> h(2 * (size_t)count);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < count; ++m) {
        long double a = -2.0L * pi * (long double)m / (long double)denom;
        h[2 * m] = (R)cosl(a);
        h[2 * m + 1] = (R)sinl(a);
    }
    (void)L;
    LSB_TRY(dev_alloc(dptr, h.size() * sizeof(R)));
    LSB_HIP(hipMemcpy(*dptr, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice));
    return LSB_OK;
}

int hor_plan_create(HorPlan* hp, int nx, int ny) {
    hp->nx = nx;
    hp->ny = ny;
    hp->lh = nx / 2;
    if (nx % 2 != 0 || !make_plan(hp->lh, &hp->px) || !make_plan(ny, &hp->py)) {
        set_error("horfft: nx must be even and nx/2, ny products of 2,3,5");
        return LSB_E_ARG;
    }
    LSB_TRY(make_twiddles<double>((void**)&hp->tw_x, hp->lh, hp->lh, hp->lh));
    LSB_TRY(make_twiddles<double>((void**)&hp->tw_y, ny, ny, ny));
    LSB_TRY(make_twiddles<double>((void**)&hp->tw_r, nx, nx, hp->lh));
    LSB_TRY(make_twiddles<float>((void**)&hp->tw_x32, hp->lh, hp->lh, hp->lh));
    LSB_TRY(make_twiddles<float>((void**)&hp->tw_y32, ny, ny, ny));
    LSB_TRY(make_twiddles<float>((void**)&hp->tw_r32, nx, nx, hp->lh));
    return LSB_OK;
}

void hor_plan_destroy(HorPlan* hp) {
    void* ps[6] = {hp->tw_x, hp->tw_y, hp->tw_r, hp->tw_x32, hp->tw_y32, hp->tw_r32};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    hp->tw_x = hp->tw_y = hp->tw_r = nullptr;
    hp->tw_x32 = hp->tw_y32 = hp->tw_r32 = nullptr;
}

// ------------------------------------------------------- small DFT kernels
template <class R, int S>
__device__ __forceinline__ cpx<R> mul_si(cpx<R> a) {  // (S*i) * a
    if (S < 0) return {a.im, -a.re};
    return {-a.im, a.re};
}

template <class R, int S>
__device__ __forceinline__ void dft2(cpx<R>* v) {
    cpx<R> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <class R, int S>
__device__ __forceinline__ void dft4(cpx<R>& a0, cpx<R>& a1, cpx<R>& a2, cpx<R>& a3) {
    cpx<R> s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = csub(a1, a3);
    cpx<R> w = mul_si<R, S>(d13);
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, w);
    a3 = csub(d02, w);
}

template <class R, int S>
__device__ __forceinline__ void dft3(cpx<R>* v) {
    const R h = (R)0.86602540378443864676372317075293618L;  // sqrt(3)/2
    cpx<R> a0 = v[0], a1 = v[1], a2 = v[2];
    cpx<R> t = cadd(a1, a2);
    cpx<R> m = {a0.re - (R)0.5 * t.re, a0.im - (R)0.5 * t.im};
    cpx<R> d = csub(a1, a2);
    cpx<R> w = mul_si<R, S>(cpx<R>{h * d.re, h * d.im});
    v[0] = cadd(a0, t);
    v[1] = cadd(m, w);
    v[2] = csub(m, w);
}

template <class R, int S>
__device__ __forceinline__ void dft5(cpx<R>* v) {
    const R c1 = (R)0.30901699437494742410229341718281906L;   // cos(2pi/5)
    const R c2 = (R)-0.80901699437494742410229341718281906L;  // cos(4pi/5)
    const R s1 = (R)0.95105651629515357211643933337938214L;   // sin(2pi/5)
    const R s2 = (R)0.58778525229247312916870595463907277L;   // sin(4pi/5)
    cpx<R> a0 = v[0];
    cpx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    cpx<R> b1 = {a0.re + c1 * t1.re + c2 * t2.re, a0.im + c1 * t1.im + c2 * t2.im};
    cpx<R> b2 = {a0.re + c2 * t1.re + c1 * t2.re, a0.im + c2 * t1.im + c1 * t2.im};
    cpx<R> e1 = mul_si<R, S>(cpx<R>{s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im});
    cpx<R> e2 = mul_si<R, S>(cpx<R>{s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im});
    v[0] = cadd(a0, cadd(t1, t2));
    v[1] = cadd(b1, e1);
    v[4] = csub(b1, e1);
    v[2] = cadd(b2, e2);
    v[3] = csub(b2, e2);
}

template <class R, int S>
__device__ __forceinline__ void dft8(cpx<R>* v) {
    const R r = (R)0.70710678118654752440084436210484904L;
    cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<R, S>(e0, e1, e2, e3);
    dft4<R, S>(o0, o1, o2, o3);
    // w^1 = r(1 + S i), w^2 = S i, w^3 = r(-1 + S i)
    cpx<R> w1 = {r * (o1.re - (R)S * o1.im), r * (o1.im + (R)S * o1.re)};
    cpx<R> w2 = mul_si<R, S>(o2);
    cpx<R> w3 = {r * (-o3.re - (R)S * o3.im), r * (-o3.im + (R)S * o3.re)};
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1);
    v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2);
    v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3);
    v[7] = csub(e3, w3);
}

template <class R, int S, int RAD>
__device__ __forceinline__ void dft_small(cpx<R>* v) {
    if (RAD == 2) dft2<R, S>(v);
    else if (RAD == 3) dft3<R, S>(v);
    else if (RAD == 4) dft4<R, S>(v[0], v[1], v[2], v[3]);
    else if (RAD == 5) dft5<R, S>(v);
    else dft8<R, S>(v);
}

// One Stockham stage of compile-time radix RAD over `nlines` lines of length L
// held in LDS (in -> out).  The butterfly operands stay in registers.
template <class R, int S, int RAD>
__device__ __forceinline__ void stockham_stage(const cpx<R>* __restrict__ in, cpx<R>* __restrict__ out, int L, int Ns,
                                               int nlines, int lstride, const cpx<R>* __restrict__ tw) {
    const int nb = L / RAD;
    const int twstep = L / (Ns * RAD);
    const int total = nlines * nb;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int line = t / nb;
        const int j = t - line * nb;
        const int jm = j % Ns;
        const cpx<R>* src = in + line * lstride;
        cpx<R>* dst = out + line * lstride;
        cpx<R> v[RAD];
#pragma unroll
        for (int r = 0; r < RAD; ++r) v[r] = src[j + r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                cpx<R> w = tw[jm * r * twstep];
                if (S > 0) w.im = -w.im;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R, S, RAD>(v);
        const int base = (j - jm) * RAD + jm;
#pragma unroll
        for (int r = 0; r < RAD; ++r) dst[base + r * Ns] = v[r];
    }
}

// Mixed-radix Stockham sequence over `nlines` lines of length L held in LDS.
// Twiddle table tw[m] = exp(-2 pi i m / L); S = -1 forward, +1 backward.
// Returns the buffer holding the result (b0 or b1).
template <class R, int S>
__device__ cpx<R>* fft_lines(cpx<R>* b0, cpx<R>* b1, const FftPlan1D& plan, int nlines, int lstride,
                             const cpx<R>* __restrict__ tw) {
    const int L = plan.L;
    cpx<R>* in = b0;
    cpx<R>* out = b1;
    int Ns = 1;
    for (int s = 0; s < plan.nstage; ++s) {
        const int rad = plan.radix[s];
        switch (rad) {
            case 2: stockham_stage<R, S, 2>(in, out, L, Ns, nlines, lstride, tw); break;
            case 3: stockham_stage<R, S, 3>(in, out, L, Ns, nlines, lstride, tw); break;
            case 4: stockham_stage<R, S, 4>(in, out, L, Ns, nlines, lstride, tw); break;
            case 5: stockham_stage<R, S, 5>(in, out, L, Ns, nlines, lstride, tw); break;
            default: stockham_stage<R, S, 8>(in, out, L, Ns, nlines, lstride, tw); break;
        }
        __syncthreads();
        cpx<R>* tmp = in;
        in = out;
        out = tmp;
        Ns *= rad;
    }
    return in;
}

// ------------------------------------------------------------- X passes
// Forward: rows of real input (stride nxpp) -> spectral rows (stride nx).
// Row index r in [0, nf*nrows); nrows = nplanes*ny per field.
template <class R>
__global__ void __launch_bounds__(256) k_xfwd(FieldList<R> in, FieldList<R> out, int nrows, int nf, int nx, int nxpp,
                                              FftPlan1D plan, const cpx<R>* __restrict__ tw,
                                              const cpx<R>* __restrict__ twr, int lines) {
    extern __shared__ unsigned char smem_raw[];
    const int lh = nx / 2;
    cpx<R>* b0 = reinterpret_cast<cpx<R>*>(smem_raw);
    cpx<R>* b1 = b0 + (size_t)lines * lh;
    __shared__ const R* rin[16];
    __shared__ R* rout[16];
    const long row0 = (long)blockIdx.x * lines;
    const long total_rows = (long)nrows * nf;
    if ((int)threadIdx.x < lines) {
        const long r = row0 + threadIdx.x;
        if (r < total_rows) {
            const int f = (int)(r / nrows);
            const long rr = r - (long)f * nrows;
            rin[threadIdx.x] = in.p[f] + rr * nxpp;
            rout[threadIdx.x] = out.p[f] + rr * nx;
        } else {
            rin[threadIdx.x] = nullptr;
            rout[threadIdx.x] = nullptr;
        }
    }
    __syncthreads();
    // load
    for (int t = threadIdx.x; t < lines * nx; t += blockDim.x) {
        const int l = t / nx, i = t - l * nx;
        const R* src = rin[l];
        reinterpret_cast<R*>(b0)[l * nx + i] = src ? src[i] : (R)0;
    }
    __syncthreads();
    cpx<R>* z = fft_lines<R, -1>(b0, b1, plan, lines, lh, tw);
    // split: X_k = 1/2 (Z_k + conj Z_{lh-k}) - i/2 W^k (Z_k - conj Z_{lh-k})
    for (int t = threadIdx.x; t < lines * lh; t += blockDim.x) {
        const int l = t / lh, k = t - l * lh;
        R* dstrow = rout[l];
        if (!dstrow) continue;
        const cpx<R> zk = z[l * lh + k];
        const cpx<R> zc = cconj(z[l * lh + ((k == 0) ? 0 : lh - k)]);
        const cpx<R> a = {(R)0.5 * (zk.re + zc.re), (R)0.5 * (zk.im + zc.im)};
        const cpx<R> d = {(R)0.5 * (zk.re - zc.re), (R)0.5 * (zk.im - zc.im)};
        const cpx<R> wd = cmul(twr[k], d);
        // -i * wd = (wd.im, -wd.re)
        const cpx<R> x = {a.re + wd.im, a.im - wd.re};
        reinterpret_cast<cpx<R>*>(dstrow)[k] = x;
    }
}

// Inverse: in place on physical rows (stride nxpp) holding nx/2 complex
// coefficients; produces nx reals and zero pads.
template <class R>
__global__ void __launch_bounds__(256) k_xinv(FieldList<R> io, int nrows, int nf, int nx, int nxpp, FftPlan1D plan,
                                              const cpx<R>* __restrict__ tw, const cpx<R>* __restrict__ twr,
                                              int lines) {
    extern __shared__ unsigned char smem_raw[];
    const int lh = nx / 2;
    cpx<R>* b0 = reinterpret_cast<cpx<R>*>(smem_raw);
    cpx<R>* b1 = b0 + (size_t)lines * lh;
    __shared__ R* rio[16];
    const long row0 = (long)blockIdx.x * lines;
    const long total_rows = (long)nrows * nf;
    if ((int)threadIdx.x < lines) {
        const long r = row0 + threadIdx.x;
        if (r < total_rows) {
            const int f = (int)(r / nrows);
            rio[threadIdx.x] = io.p[f] + (r - (long)f * nrows) * nxpp;
        } else {
            rio[threadIdx.x] = nullptr;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < lines * nx; t += blockDim.x) {
        const int l = t / nx, i = t - l * nx;
        const R* src = rio[l];
        reinterpret_cast<R*>(b1)[l * nx + i] = src ? src[i] : (R)0;
    }
    __syncthreads();
    // Z'_k = (X_k + conj X_{lh-k}) + i W^-k (X_k - conj X_{lh-k}); X_lh = 0, Im X_0 ignored
    for (int t = threadIdx.x; t < lines * lh; t += blockDim.x) {
        const int l = t / lh, k = t - l * lh;
        cpx<R> xk = b1[(size_t)l * lh + k];
        cpx<R> xc;
        if (k == 0) {
            xk.im = 0;
            xc = {0, 0};
        } else {
            xc = cconj(b1[(size_t)l * lh + (lh - k)]);
        }
        const cpx<R> s = cadd(xk, xc);
        const cpx<R> d = csub(xk, xc);
        const cpx<R> wd = cmul(cconj(twr[k]), d);
        // i * wd = (-wd.im, wd.re)
        b0[(size_t)l * lh + k] = {s.re - wd.im, s.im + wd.re};
    }
    __syncthreads();
    cpx<R>* y = fft_lines<R, 1>(b0, b1, plan, lines, lh, tw);
    for (int t = threadIdx.x; t < lines * nxpp; t += blockDim.x) {
        const int l = t / nxpp, i = t - l * nxpp;
        R* dst = rio[l];
        if (!dst) continue;
        dst[i] = (i < nx) ? reinterpret_cast<R*>(y)[l * nx + i] : (R)0;
    }
}

// ------------------------------------------------------------- Y passes
// Forward, in place on spectral planes: tiles of `cols` kx columns x ny rows.
template <class R>
__global__ void __launch_bounds__(256) k_yfwd(FieldList<R> io, int nplanes, int nx, int ny, FftPlan1D plan,
                                              const cpx<R>* __restrict__ tw, int cols, R scale) {
    extern __shared__ unsigned char smem_raw[];
    const int lh = nx / 2;
    cpx<R>* b0 = reinterpret_cast<cpx<R>*>(smem_raw);
    cpx<R>* b1 = b0 + (size_t)cols * ny;
    const int ntile = (lh + cols - 1) / cols;
    const int tile = blockIdx.x % ntile;
    const long plane = blockIdx.x / ntile;  // over nf * nplanes
    const int f = (int)(plane / nplanes);
    const long pk = plane - (long)f * nplanes;
    cpx<R>* base = reinterpret_cast<cpx<R>*>(io.p[f] + pk * (long)nx * ny);
    const int m0 = tile * cols;
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        const int j = t / cols, cc = t - j * cols;
        const int m = m0 + cc;
        b0[(size_t)cc * ny + j] = (m < lh) ? base[(size_t)j * lh + m] : cpx<R>{0, 0};
    }
    __syncthreads();
    cpx<R>* y = fft_lines<R, -1>(b0, b1, plan, cols, ny, tw);
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        const int j = t / cols, cc = t - j * cols;
        const int m = m0 + cc;
        if (m < lh) {
            cpx<R> v = y[(size_t)cc * ny + j];
            base[(size_t)j * lh + m] = {v.re * scale, v.im * scale};
        }
    }
}

// Inverse: spectral planes (read) -> physical buffer rows (complex stride nxpp/2).
template <class R>
__global__ void __launch_bounds__(256) k_yinv(FieldList<R> in, FieldList<R> out, int nplanes, int nx, int ny,
                                              int nxpp, FftPlan1D plan, const cpx<R>* __restrict__ tw, int cols) {
    extern __shared__ unsigned char smem_raw[];
    const int lh = nx / 2;
    const int ph = nxpp / 2;
    cpx<R>* b0 = reinterpret_cast<cpx<R>*>(smem_raw);
    cpx<R>* b1 = b0 + (size_t)cols * ny;
    const int ntile = (lh + cols - 1) / cols;
    const int tile = blockIdx.x % ntile;
    const long plane = blockIdx.x / ntile;
    const int f = (int)(plane / nplanes);
    const long pk = plane - (long)f * nplanes;
    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(in.p[f] + pk * (long)nx * ny);
    cpx<R>* dst = reinterpret_cast<cpx<R>*>(out.p[f] + pk * (long)nxpp * ny);
    const int m0 = tile * cols;
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        const int j = t / cols, cc = t - j * cols;
        const int m = m0 + cc;
        b0[(size_t)cc * ny + j] = (m < lh) ? src[(size_t)j * lh + m] : cpx<R>{0, 0};
    }
    __syncthreads();
    cpx<R>* y = fft_lines<R, 1>(b0, b1, plan, cols, ny, tw);
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        const int j = t / cols, cc = t - j * cols;
        const int m = m0 + cc;
        if (m < lh) dst[(size_t)j * ph + m] = y[(size_t)cc * ny + j];
    }
}

// ------------------------------------------------------------- launchers
static int lines_for(int len, size_t elem) {
    int l = (int)(32768 / ((size_t)len * elem));
    if (l < 1) l = 1;
    if (l > 16) l = 16;
    return l;
}

static double fft_flops(const HorPlan& hp, int nf, int nplanes) {
    const double N = (double)hp.nx * hp.ny;
    return (double)nf * nplanes * 2.5 * N * std::log2(N);
}
// algorithmic bytes of nf plane-batched 2-D transforms: read one layout, write the other
static double fft_bytes(const HorPlan& hp, int nf, int nplanes, size_t phys_elem, size_t spec_elem) {
    const double planes = (double)nf * nplanes;
    return planes * ((double)(hp.nx + 2) * hp.ny * phys_elem + (double)hp.nx * hp.ny * spec_elem);
}

int horfft_fwd(lsb_ctx* c, const HorPlan& hp, const double* const* phys, double* const* spec, int nf, int nplanes,
               hipStream_t s) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("horfft: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    FieldList<double> in{}, out{};
    for (int f = 0; f < nf; ++f) {
        in.p[f] = const_cast<double*>(phys[f]);
        out.p[f] = spec[f];
    }
    int tok = -1;
    prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, nf, nplanes), fft_bytes(hp, nf, nplanes, 8, 8));
    const int nrows = nplanes * ny;
    if (fast_fft_len(lh)) {
        LSB_TRY((fast_xfwd<double, double, double>(in, out, nrows, nf, lh, nxpp, nx, (const cpx<double>*)hp.tw_x,
                                                   (const cpx<double>*)hp.tw_r, s)));
    } else {
        int lines = lines_for(lh, sizeof(cpx<double>));
        size_t shm = 2 * (size_t)lines * lh * sizeof(cpx<double>);
        long nblk = ((long)nrows * nf + lines - 1) / lines;
        hipLaunchKernelGGL(k_xfwd<double>, dim3((unsigned)nblk), dim3(256), shm, s, in, out, nrows, nf, nx, nxpp,
                           hp.px, (const cpx<double>*)hp.tw_x, (const cpx<double>*)hp.tw_r, lines);
        LSB_HIP(hipGetLastError());
    }
    const double scale = 1.0 / (double)(nx * ny);
    if (fast_fft_len(ny)) {
        LSB_TRY((fast_yfwd<double, double, double>(out, out, nplanes, nf, lh, ny, (const cpx<double>*)hp.tw_y, scale,
                                                   s, c->tune[TUNE_YFWD_F64])));
    } else {
        int cols = lines_for(ny, sizeof(cpx<double>));
        size_t shm = 2 * (size_t)cols * ny * sizeof(cpx<double>);
        const int ntile = (lh + cols - 1) / cols;
        long nblk = (long)ntile * nplanes * nf;
        hipLaunchKernelGGL(k_yfwd<double>, dim3((unsigned)nblk), dim3(256), shm, s, out, nplanes, nx, ny, hp.py,
                           (const cpx<double>*)hp.tw_y, cols, scale);
        LSB_HIP(hipGetLastError());
    }
    prof_mark_end(c, tok, s);
    return LSB_OK;
}

bool horfft_has_ops(const HorPlan& hp) { return fast_fft_len(hp.ny) && fast_fft_len(hp.lh); }

int horfft_inv_ops(lsb_ctx* c, const HorPlan& hp, const double* const* spec, double* const* phys, int nf,
                   int nplanes, const int* ops, hipStream_t s, double* scratch) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("horfft: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    FieldList<double> in{}, out{};
    FftOps fo{};
    bool any_op = false;
    for (int f = 0; f < nf; ++f) {
        in.p[f] = const_cast<double*>(spec[f]);
        out.p[f] = phys[f];
        fo.op[f] = ops ? ops[f] : FOP_NONE;
        any_op = any_op || fo.op[f] != FOP_NONE;
    }
    if (any_op && !horfft_has_ops(hp)) {
        set_error("horfft: derivative-on-load needs power-of-two nx/2 and ny in 16..256");
        return LSB_E_UNSUPPORTED;
    }
    int tok = -1;
    prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, nf, nplanes), fft_bytes(hp, nf, nplanes, 8, 8));
    const int nrows = nplanes * ny;
    if (horfft_has_ops(hp)) {
        // y pass into a compact scratch (rows of lh complex, 128-byte aligned), then the
        // x pass out of place into the padded physical rows
        void* scr = scratch;
        const size_t per = (size_t)nplanes * ny * nx;
        if (!scr) LSB_TRY(stage_get(c, 8, per * nf * sizeof(double), &scr));
        FieldList<double> mid{};
        for (int f = 0; f < nf; ++f) mid.p[f] = (double*)scr + per * f;
        LSB_TRY((fast_yinv<double, double, double>(in, mid, nplanes, nf, lh, ny, lh, (const cpx<double>*)hp.tw_y, fo,
                                                   c->d_wavx, c->d_wavy, s)));
        LSB_TRY((fast_xinv<double, double, double>(mid, out, nrows, nf, lh, nx, nxpp, (const cpx<double>*)hp.tw_x,
                                                   (const cpx<double>*)hp.tw_r, s)));
        prof_mark_end(c, tok, s);
        return LSB_OK;
    }
    if (fast_fft_len(ny)) {
        LSB_TRY((fast_yinv<double, double, double>(in, out, nplanes, nf, lh, ny, nxpp / 2, (const cpx<double>*)hp.tw_y,
                                                   fo, c->d_wavx, c->d_wavy, s)));
    } else {
        int cols = lines_for(ny, sizeof(cpx<double>));
        size_t shm = 2 * (size_t)cols * ny * sizeof(cpx<double>);
        const int ntile = (lh + cols - 1) / cols;
        long nblk = (long)ntile * nplanes * nf;
        hipLaunchKernelGGL(k_yinv<double>, dim3((unsigned)nblk), dim3(256), shm, s, in, out, nplanes, nx, ny, nxpp,
                           hp.py, (const cpx<double>*)hp.tw_y, cols);
        LSB_HIP(hipGetLastError());
    }
    if (fast_fft_len(lh)) {
        LSB_TRY((fast_xinv<double, double, double>(out, out, nrows, nf, lh, nxpp, nxpp, (const cpx<double>*)hp.tw_x,
                                                   (const cpx<double>*)hp.tw_r, s)));
    } else {
        int lines = lines_for(lh, sizeof(cpx<double>));
        size_t shm = 2 * (size_t)lines * lh * sizeof(cpx<double>);
        long nblk = ((long)nrows * nf + lines - 1) / lines;
        hipLaunchKernelGGL(k_xinv<double>, dim3((unsigned)nblk), dim3(256), shm, s, out, nrows, nf, nx, nxpp, hp.px,
                           (const cpx<double>*)hp.tw_x, (const cpx<double>*)hp.tw_r, lines);
        LSB_HIP(hipGetLastError());
    }
    prof_mark_end(c, tok, s);
    return LSB_OK;
}

int horfft_inv(lsb_ctx* c, const HorPlan& hp, const double* const* spec, double* const* phys, int nf, int nplanes,
               hipStream_t s) {
    return horfft_inv_ops(c, hp, spec, phys, nf, nplanes, nullptr, s);
}

// ------------------------------------------- mixed form, generic lengths
// Lengths that are not register-resident (nx/2 or ny with a factor 3 or 5, or outside
// 16..256) run the mixed form's transforms on the Stockham kernels in FP32.

// u, du/dx, du/dy of 3 FP64 spectra rounded to 9 FP32 spectra; the derivatives are
// taken in FP64 first (derivatives.f90:261-303: i kx f, i ky f).  One thread per
// (field, pair); pairs are (plane, ky, kx) with kx fastest.
__global__ void k_grad_spec32(FieldList<double> in, FieldList<float> out, long npairs, int lh, int ny,
                              const double* __restrict__ wavx, const double* __restrict__ wavy) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * npairs) return;
    const int l = (int)(t / npairs);
    const long pr = t - (long)l * npairs;
    const int m = (int)(pr % lh);
    const int j = (int)((pr / lh) % ny);
    const double2 v = reinterpret_cast<const double2*>(in.p[l])[pr];
    const double wx = wavx[2 * m], wy = wavy[j];
    reinterpret_cast<float2*>(out.p[3 * l])[pr] = make_float2((float)v.x, (float)v.y);
    reinterpret_cast<float2*>(out.p[3 * l + 1])[pr] = make_float2((float)(-wx * v.y), (float)(wx * v.x));
    reinterpret_cast<float2*>(out.p[3 * l + 2])[pr] = make_float2((float)(-wy * v.y), (float)(wy * v.x));
}

// FP32 spectra widened to FP64, n values per field
__global__ void k_widen_f32(FieldList<float> in, FieldList<double> out, long n, int nf) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nf) return;
    const int f = (int)(t / n);
    const long i = t - (long)f * n;
    out.p[f][i] = (double)in.p[f][i];
}

// inverse y pass then x pass of nf FP32 spectra into padded FP32 physical rows
static int generic_inv32(const HorPlan& hp, const FieldList<float>& spec, const FieldList<float>& phys, int nf,
                         int nplanes, hipStream_t s) {
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    {
        const int cols = lines_for(ny, sizeof(cpx<float>));
        const size_t shm = 2 * (size_t)cols * ny * sizeof(cpx<float>);
        const long nblk = (long)((lh + cols - 1) / cols) * nplanes * nf;
        hipLaunchKernelGGL(k_yinv<float>, dim3((unsigned)nblk), dim3(256), shm, s, spec, phys, nplanes, nx, ny, nxpp,
                           hp.py, (const cpx<float>*)hp.tw_y32, cols);
        LSB_HIP(hipGetLastError());
    }
    {
        const int lines = lines_for(lh, sizeof(cpx<float>));
        const size_t shm = 2 * (size_t)lines * lh * sizeof(cpx<float>);
        const long nblk = ((long)nplanes * ny * nf + lines - 1) / lines;
        hipLaunchKernelGGL(k_xinv<float>, dim3((unsigned)nblk), dim3(256), shm, s, phys, nplanes * ny, nf, nx, nxpp,
                           hp.px, (const cpx<float>*)hp.tw_x32, (const cpx<float>*)hp.tw_r32, lines);
        LSB_HIP(hipGetLastError());
    }
    return LSB_OK;
}

// forward x pass then y pass of nf FP32 physical fields into FP32 spectra
static int generic_fwd32(const HorPlan& hp, const FieldList<float>& phys, const FieldList<float>& spec, int nf,
                         int nplanes, hipStream_t s) {
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    {
        const int lines = lines_for(lh, sizeof(cpx<float>));
        const size_t shm = 2 * (size_t)lines * lh * sizeof(cpx<float>);
        const long nblk = ((long)nplanes * ny * nf + lines - 1) / lines;
        hipLaunchKernelGGL(k_xfwd<float>, dim3((unsigned)nblk), dim3(256), shm, s, phys, spec, nplanes * ny, nf, nx,
                           nxpp, hp.px, (const cpx<float>*)hp.tw_x32, (const cpx<float>*)hp.tw_r32, lines);
        LSB_HIP(hipGetLastError());
    }
    {
        const int cols = lines_for(ny, sizeof(cpx<float>));
        const size_t shm = 2 * (size_t)cols * ny * sizeof(cpx<float>);
        const long nblk = (long)((lh + cols - 1) / cols) * nplanes * nf;
        hipLaunchKernelGGL(k_yfwd<float>, dim3((unsigned)nblk), dim3(256), shm, s, spec, nplanes, nx, ny, hp.py,
                           (const cpx<float>*)hp.tw_y32, cols, (float)(1.0 / ((double)nx * ny)));
        LSB_HIP(hipGetLastError());
    }
    return LSB_OK;
}

// the gradient transform of the mixed form at generic lengths: 9 FP32 spectra in a
// staging buffer (k_grad_spec32), then 9 FP32 inverse transforms
static int inv_grad32_generic(lsb_ctx* c, const HorPlan& hp, const double* const* uif, float* const* u,
                              float* const* dudx, float* const* dudy, int nplanes, hipStream_t s) {
    const long per = (long)nplanes * hp.ny * hp.nx;  // floats of one spectrum
    const long npairs = per / 2;
    void* scr = nullptr;
    LSB_TRY(stage_get(c, 8, (size_t)per * 9 * sizeof(float), &scr));
    FieldList<double> in{};
    FieldList<float> mid{}, out{};
    for (int l = 0; l < 3; ++l) {
        in.p[l] = const_cast<double*>(uif[l]);
        out.p[3 * l] = u[l];
        out.p[3 * l + 1] = dudx[l];
        out.p[3 * l + 2] = dudy[l];
    }
    for (int f = 0; f < 9; ++f) mid.p[f] = (float*)scr + per * f;
    int tok = -1;
    {
        const double planes = (double)nplanes;
        const double bytes = planes * (3.0 * hp.nx * hp.ny * 8 + 9.0 * (hp.nx + 2) * hp.ny * sizeof(float));
        prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, 9, nplanes), bytes);
    }
    hipLaunchKernelGGL(k_grad_spec32, dim3((unsigned)((3 * npairs + 255) / 256)), dim3(256), 0, s, in, mid, npairs,
                       hp.lh, hp.ny, (const double*)c->d_wavx, (const double*)c->d_wavy);
    LSB_HIP(hipGetLastError());
    LSB_TRY(generic_inv32(hp, mid, out, 9, nplanes, s));
    prof_mark_end(c, tok, s);
    return LSB_OK;
}

// u, du/dx, du/dy of 3 velocity spectra in one pass over each spectrum: the inverse
// y pass produces the three transforms from one load (fast_yinv3), the x pass
// writes the padded physical rows.  RT is the physical storage type and the
// arithmetic type (double: FP64 form; float: mixed form).  At generic lengths only
// the mixed form is served here (the FP64 form takes spectral derivatives and the
// plain inverse transform), and the spectral filter is not applied on load.
template <class RT>
int horfft_inv_grad(lsb_ctx* c, const HorPlan& hp, const double* const* uif, RT* const* u, RT* const* dudx,
                    RT* const* dudy, int nplanes, const double* ftab, hipStream_t s) {
    if (!horfft_has_ops(hp)) {
        if constexpr (sizeof(RT) == 4) {
            if (!ftab) return inv_grad32_generic(c, hp, uif, u, dudx, dudy, nplanes, s);
        }
        set_error("horfft: the fused FP64 gradient transform and the filter on load need power-of-two nx/2 and ny "
                  "in 16..256");
        return LSB_E_UNSUPPORTED;
    }
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    void* scr = nullptr;
    const size_t per = (size_t)nplanes * ny * nx;
    LSB_TRY(stage_get(c, 8, per * 9 * sizeof(RT), &scr));
    FieldList<double> in{};
    FieldList<RT> mid{}, out{};
    for (int l = 0; l < 3; ++l) {
        in.p[l] = const_cast<double*>(uif[l]);
        out.p[3 * l] = u[l];
        out.p[3 * l + 1] = dudx[l];
        out.p[3 * l + 2] = dudy[l];
    }
    for (int f = 0; f < 9; ++f) mid.p[f] = (RT*)scr + per * f;
    int tok = -1;
    {
        const double planes = (double)nplanes;
        const double bytes = planes * (3.0 * hp.nx * hp.ny * 8 + 9.0 * (hp.nx + 2) * hp.ny * sizeof(RT));
        prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, 9, nplanes), bytes);
    }
    const cpx<RT>* twy = (const cpx<RT>*)(sizeof(RT) == 8 ? (void*)hp.tw_y : (void*)hp.tw_y32);
    const cpx<RT>* twx = (const cpx<RT>*)(sizeof(RT) == 8 ? (void*)hp.tw_x : (void*)hp.tw_x32);
    const cpx<RT>* twr = (const cpx<RT>*)(sizeof(RT) == 8 ? (void*)hp.tw_r : (void*)hp.tw_r32);
    LSB_TRY((fast_yinv3<RT, double, RT>(in, mid, nplanes, 3, lh, ny, lh, twy, c->d_wavx, c->d_wavy, ftab, s,
                                        c->tune[sizeof(RT) == 8 ? TUNE_YINV3_F64 : TUNE_YINV3_F32])));
    LSB_TRY((fast_xinv<RT, RT, RT>(mid, out, nplanes * ny, 9, lh, nx, nxpp, twx, twr, s)));
    prof_mark_end(c, tok, s);
    return LSB_OK;
}
template int horfft_inv_grad<double>(lsb_ctx*, const HorPlan&, const double* const*, double* const*, double* const*,
                                     double* const*, int, const double*, hipStream_t);
template int horfft_inv_grad<float>(lsb_ctx*, const HorPlan&, const double* const*, float* const*, float* const*,
                                    float* const*, int, const double*, hipStream_t);

// Mixed form: FP32 physical fields, FP64 spectra, FP32 arithmetic.
template <class TO>
int horfft_fwd_mixed(lsb_ctx* c, const HorPlan& hp, const float* const* phys, TO* const* spec, int nf, int nplanes,
                     hipStream_t s) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("horfft (mixed): 1..16 fields per launch");
        return LSB_E_ARG;
    }
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    void* scr = nullptr;
    const size_t per = (size_t)nplanes * ny * nx;
    if (!horfft_has_ops(hp)) {
        // generic lengths: Stockham passes in FP32 (into staging first when the spectra are FP64)
        FieldList<float> in{}, mid{};
        FieldList<double> wide{};
        if (sizeof(TO) == 8) LSB_TRY(stage_get(c, 8, per * nf * sizeof(float), &scr));
        for (int f = 0; f < nf; ++f) {
            in.p[f] = const_cast<float*>(phys[f]);
            mid.p[f] = sizeof(TO) == 8 ? (float*)scr + per * f : reinterpret_cast<float*>(spec[f]);
            wide.p[f] = reinterpret_cast<double*>(spec[f]);
        }
        int tok = -1;
        prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, nf, nplanes), fft_bytes(hp, nf, nplanes, 4, sizeof(TO)));
        LSB_TRY(generic_fwd32(hp, in, mid, nf, nplanes, s));
        if (sizeof(TO) == 8) {
            hipLaunchKernelGGL(k_widen_f32, dim3((unsigned)(((long)per * nf + 255) / 256)), dim3(256), 0, s, mid, wide,
                               (long)per, nf);
            LSB_HIP(hipGetLastError());
        }
        prof_mark_end(c, tok, s);
        return LSB_OK;
    }
    LSB_TRY(stage_get(c, 8, per * nf * sizeof(float), &scr));
    FieldList<float> in{}, mid{};
    FieldList<TO> out{};
    for (int f = 0; f < nf; ++f) {
        in.p[f] = const_cast<float*>(phys[f]);
        mid.p[f] = (float*)scr + per * f;
        out.p[f] = spec[f];
    }
    int tok = -1;
    prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, nf, nplanes), fft_bytes(hp, nf, nplanes, 4, sizeof(TO)));
    const int nrows = nplanes * ny;
    LSB_TRY((fast_xfwd<float, float, float>(in, mid, nrows, nf, lh, nxpp, nx, (const cpx<float>*)hp.tw_x32,
                                            (const cpx<float>*)hp.tw_r32, s)));
    LSB_TRY((fast_yfwd<float, float, TO>(mid, out, nplanes, nf, lh, ny, (const cpx<float>*)hp.tw_y32,
                                             1.0f / (float)(nx * ny), s, c->tune[TUNE_YFWD_F32])));
    prof_mark_end(c, tok, s);
    return LSB_OK;
}
template int horfft_fwd_mixed<double>(lsb_ctx*, const HorPlan&, const float* const*, double* const*, int, int,
                                      hipStream_t);
template int horfft_fwd_mixed<float>(lsb_ctx*, const HorPlan&, const float* const*, float* const*, int, int,
                                     hipStream_t);

// ------------------------------------------------ y-pass tile autotuning
// Times the kx-column tiles of the y passes the step uses (forward y pass, the gradient
// transform's inverse y pass, the mixed form's forward y pass fused with the right-hand
// side) on this context's grid, FP64 and FP32 arithmetic, and keeps the fastest.  work:
// zeroed scratch of at least 24 physical fields of doubles.
int fft_tune_y(lsb_ctx* c, void* work, hipStream_t s) {
    const HorPlan& hp = c->hp;
    if (!horfft_has_ops(hp)) return LSB_OK;
    const int np = c->nzp, lh = hp.lh, ny = hp.ny;
    const size_t nsp = (size_t)np * ny * hp.nx;  // reals of one spectral field
    char* w = static_cast<char*>(work);
    auto f64 = [&](size_t i) { return reinterpret_cast<double*>(w) + i * nsp; };
    auto f32 = [&](size_t i) { return reinterpret_cast<float*>(w) + i * nsp; };
    hipEvent_t e0, e1;
    LSB_HIP(hipEventCreate(&e0));
    LSB_HIP(hipEventCreate(&e1));
    const double scale = 1.0 / ((double)hp.nx * ny);
    const bool verbose = std::getenv("LSB_TUNE_VERBOSE") != nullptr;
    int rc = LSB_OK;
    // one warm-up launch, then the mean of three
    auto pick = [&](int slot, const char* name, const int* cand, int ncand, auto&& launch) {
        float best = 1e30f;
        int bc = -1;
        for (int i = 0; i < ncand && rc == LSB_OK; ++i) {
            rc = launch(cand[i]);
            if (rc != LSB_OK) break;
            if (hipEventRecord(e0, s) != hipSuccess) rc = LSB_E_HIP;
            for (int r = 0; r < 3 && rc == LSB_OK; ++r) rc = launch(cand[i]);
            if (rc == LSB_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                rc = LSB_E_HIP;
            float ms = 0.f;
            if (rc == LSB_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = LSB_E_HIP;
            if (rc != LSB_OK) break;
            if (verbose) std::fprintf(stderr, "lsb_autotune %s: %d columns %.4f ms\n", name, cand[i], ms / 3.0f);
            if (ms < best) {
                best = ms;
                bc = cand[i];
            }
        }
        if (rc == LSB_OK && bc > 0) {
            c->tune[slot] = bc;
            c->tune_ms[slot] = best / 3.0f;
        }
    };
    const int c64[2] = {8, 16}, c32[3] = {8, 16, 32};
    {   // forward y pass: 3 fields
        FieldList<double> in{}, out{};
        FieldList<float> in32{}, out32{};
        for (int f = 0; f < 3; ++f) {
            in.p[f] = f64(f);
            out.p[f] = f64(3 + f);
            in32.p[f] = f32(f);
            out32.p[f] = f32(3 + f);
        }
        pick(TUNE_YFWD_F64, "yfwd f64", c64, 2, [&](int cc) {
            return fast_yfwd<double, double, double>(in, out, np, 3, lh, ny, (const cpx<double>*)hp.tw_y, scale, s,
                                                     cc);
        });
        pick(TUNE_YFWD_F32, "yfwd f32", c32, 3, [&](int cc) {
            return fast_yfwd<float, float, float>(in32, out32, np, 3, lh, ny, (const cpx<float>*)hp.tw_y32,
                                                  (float)scale, s, cc);
        });
    }
    {   // gradient transform's inverse y pass: 3 FP64 spectra -> 9 fields
        FieldList<double> in{}, out{};
        FieldList<float> out32{};
        for (int f = 0; f < 3; ++f) in.p[f] = f64(f);
        for (int f = 0; f < 9; ++f) {
            out.p[f] = f64(3 + f);
            out32.p[f] = f32(6 + f);  // after the 3 FP64 inputs (6 FP32 fields of room)
        }
        pick(TUNE_YINV3_F64, "yinv3 f64", c64, 2, [&](int cc) {
            return fast_yinv3<double, double, double>(in, out, np, 3, lh, ny, lh, (const cpx<double>*)hp.tw_y,
                                                      c->d_wavx, c->d_wavy, nullptr, s, cc);
        });
        pick(TUNE_YINV3_F32, "yinv3 f32", c32, 3, [&](int cc) {
            return fast_yinv3<float, double, float>(in, out32, np, 3, lh, ny, lh, (const cpx<float>*)hp.tw_y32,
                                                    c->d_wavx, c->d_wavy, nullptr, s, cc);
        });
    }
    {   // mixed form's forward y pass with the right-hand side: 9 FP32 pieces, FP32 N
        // history, FP64 velocity history and right-hand sides (168 bytes per spectral real)
        FieldList<float> pc{};
        YRhsArgs a{};
        for (int f = 0; f < 9; ++f) pc.p[f] = f32(f);
        for (int l = 0; l < 3; ++l) {
            a.nout[l] = f32(9 + l);
            a.nl[l] = f32(12 + l);
            a.nlm[l] = f32(15 + l);
            a.u[l] = f64(9 + l);
            a.un[l] = f64(12 + l);
            a.unm[l] = f64(15 + l);
            a.rhs[l] = f64(18 + l);
        }
        a.wavx = c->d_wavx;
        a.wavy = c->d_wavy;
        a.ubarpp = c->d_paral + 2 * c->nzp;
        a.ab0 = 3.0;
        a.ab1 = -3.0;
        a.ab2 = 1.0;
        a.bdf0 = 3.0;
        a.bdf1 = -1.5;
        a.bdf2 = 1.0 / 3.0;
        a.dt = c->dt;
        a.dtnui = 1.0;
        pick(TUNE_YRHS_F32, "yfwd+rhs f32", c32, 3, [&](int cc) {
            return fast_yfwd_rhs<float, float>(pc, np, lh, ny, (const cpx<float>*)hp.tw_y32, (float)scale, a, s, cc);
        });
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// x pass only of nf forward transforms: physical rows -> rows of lh complex (the
// layout of a spectral field before its y pass), for a fused consumer of the y pass.
template <class T>
int horfft_xfwd(lsb_ctx* c, const HorPlan& hp, const T* const* phys, T* const* rows, int nf, int nplanes,
                hipStream_t s) {
    if (nf < 1 || nf > kMaxFields || !horfft_has_ops(hp)) {
        set_error("horfft (x pass): 1..16 fields, power-of-two nx/2 and ny in 16..256");
        return LSB_E_UNSUPPORTED;
    }
    FieldList<T> in{}, out{};
    for (int f = 0; f < nf; ++f) {
        in.p[f] = const_cast<T*>(phys[f]);
        out.p[f] = rows[f];
    }
    int tok = -1;
    const double planes = (double)nf * nplanes;
    prof_mark_begin(c, PC_FFT, s, &tok, planes * 2.5 * hp.nx * hp.ny * std::log2((double)hp.nx),
                    planes * ((double)(hp.nx + 2) * hp.ny + (double)hp.nx * hp.ny) * sizeof(T));
    const cpx<T>* twx = (const cpx<T>*)(sizeof(T) == 8 ? (void*)hp.tw_x : (void*)hp.tw_x32);
    const cpx<T>* twr = (const cpx<T>*)(sizeof(T) == 8 ? (void*)hp.tw_r : (void*)hp.tw_r32);
    LSB_TRY((fast_xfwd<T, T, T>(in, out, nplanes * hp.ny, nf, hp.lh, hp.nx + 2, hp.nx, twx, twr, s)));
    prof_mark_end(c, tok, s);
    return LSB_OK;
}
template int horfft_xfwd<double>(lsb_ctx*, const HorPlan&, const double* const*, double* const*, int, int,
                                 hipStream_t);
template int horfft_xfwd<float>(lsb_ctx*, const HorPlan&, const float* const*, float* const*, int, int, hipStream_t);

// y pass of the three pieces of each N_l fused with the velocity right-hand side
// (fast_yfwd_rhs); pieces[3 l + q] hold x-transformed rows.
template <class R, class TH>
int horfft_yfwd_rhs(lsb_ctx* c, const HorPlan& hp, R* const* pieces, int nplanes, const YRhsArgs& a, hipStream_t s) {
    if (!horfft_has_ops(hp)) {
        set_error("horfft (fused y pass): power-of-two nx/2 and ny in 16..256");
        return LSB_E_UNSUPPORTED;
    }
    FieldList<R> pc{};
    for (int q = 0; q < 9; ++q) pc.p[q] = pieces[q];
    int tok = -1;
    {
        const double pairs = (double)nplanes * hp.ny * hp.lh * 3.0;
        const double per = 3.0 * 2 * sizeof(R) + 3.0 * 2 * sizeof(TH) + 4.0 * 16.0;
        prof_mark_begin(c, PC_RHS, s, &tok, 9.0 * nplanes * 2.5 * hp.lh * hp.ny * std::log2((double)hp.ny),
                        pairs * per);
    }
    const cpx<R>* twy = (const cpx<R>*)(sizeof(R) == 8 ? (void*)hp.tw_y : (void*)hp.tw_y32);
    LSB_TRY((fast_yfwd_rhs<R, TH>(pc, nplanes, hp.lh, hp.ny, twy, (R)(1.0 / ((double)hp.nx * hp.ny)), a, s,
                                  sizeof(R) == 4 ? c->tune[TUNE_YRHS_F32] : 0)));
    prof_mark_end(c, tok, s);
    return LSB_OK;
}
template int horfft_yfwd_rhs<double, double>(lsb_ctx*, const HorPlan&, double* const*, int, const YRhsArgs&,
                                             hipStream_t);
template int horfft_yfwd_rhs<float, float>(lsb_ctx*, const HorPlan&, float* const*, int, const YRhsArgs&,
                                           hipStream_t);

template <class TI>
int horfft_inv_mixed(lsb_ctx* c, const HorPlan& hp, const TI* const* spec, float* const* phys, int nf, int nplanes,
                     const int* ops, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields || !horfft_has_ops(hp)) {
        set_error("horfft (mixed): 1..16 fields, power-of-two nx/2 and ny in 16..256");
        return LSB_E_UNSUPPORTED;
    }
    const int nx = hp.nx, ny = hp.ny, lh = hp.lh, nxpp = nx + 2;
    void* scr = nullptr;
    const size_t per = (size_t)nplanes * ny * nx;
    LSB_TRY(stage_get(c, 8, per * nf * sizeof(float), &scr));
    FieldList<TI> in{};
    FieldList<float> mid{}, out{};
    FftOps fo{};
    for (int f = 0; f < nf; ++f) {
        in.p[f] = const_cast<TI*>(spec[f]);
        mid.p[f] = (float*)scr + per * f;
        out.p[f] = phys[f];
        fo.op[f] = ops ? ops[f] : FOP_NONE;
    }
    int tok = -1;
    prof_mark_begin(c, PC_FFT, s, &tok, fft_flops(hp, nf, nplanes), fft_bytes(hp, nf, nplanes, 4, sizeof(TI)));
    const int nrows = nplanes * ny;
    LSB_TRY((fast_yinv<float, TI, float>(in, mid, nplanes, nf, lh, ny, lh, (const cpx<float>*)hp.tw_y32, fo,
                                         c->d_wavx, c->d_wavy, s)));
    LSB_TRY((fast_xinv<float, float, float>(mid, out, nrows, nf, lh, nx, nxpp, (const cpx<float>*)hp.tw_x32,
                                            (const cpx<float>*)hp.tw_r32, s)));
    prof_mark_end(c, tok, s);
    return LSB_OK;
}
template int horfft_inv_mixed<double>(lsb_ctx*, const HorPlan&, const double* const*, float* const*, int, int,
                                      const int*, hipStream_t);
template int horfft_inv_mixed<float>(lsb_ctx*, const HorPlan&, const float* const*, float* const*, int, int,
                                     const int*, hipStream_t);

}  // namespace lsb
