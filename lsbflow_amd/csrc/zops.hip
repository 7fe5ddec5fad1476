// Wall-normal (Chebyshev) operators as batched dense contractions.
//
//   dfdz / d2fdz2 / dpdz  (derivatives.f90:112-188 -> chebyshev.f90:204-270):
//       out[k][p] = sum_m M[k][m] in[m][p]
//   solve_helmholtz       (helmholtz.f90:130-208), two contractions:
//       t[r][p]   = (sum_m Einv[r][m] q[m+1][p]) / (lam_r - kx^2 - ky^2 - sigma)
//       q[r+1][p] = sum_m E[r][m] t[m][p],   q[0][p] = q[nz][p] = 0
//   solve_nhom_helmholtz  (helmholtz.f90:211-358): the lifting subtraction
//       q - L(h_bot) bc_bot - L(h_top) bc_top, L(h) = h'' - (kx^2+ky^2+sigma) h, is
//       applied after the stage-1 projection as Einv q - (Einv L(h)) bc, with Einv h
//       and Einv h'' precomputed; the h*bc addition is fused into the stage-2 store.
//
// A spectral field is a row-major [nzp][P] matrix with P = nx*ny reals, so every
// contraction is M (nout x nin) times a tall skinny panel; each (Re or Im) column
// is one (kx,ky) pencil.  FP64 path on the matrix cores: v_mfma_f64_16x16x4f64
// tiles.  A block holds up to 8 waves that share 64 pencils (4 column blocks of
// 16); wave w owns at most 3 of the ceil(nout/16) row blocks, so each lane keeps
// 3 x 4 accumulator tiles (48 doubles).  Fragments come straight from memory:
// A[i][k] = M[r0+i][k0+k] on lane i + 16k (16 consecutive matrix rows, one
// 128-byte segment per k), B[k][j] = X[k0+k][p0+j] on lane j + 16k (16 consecutive
// pencils); the panel rows of k-step s+2 are in flight while step s is multiplied.
// Output lane layout (measured): lane l, element v holds D[l/16 + 4v][l%16].
// On gfx950 the FP64 matrix rate equals the FP64 vector rate (78.6 TFLOP/s); the
// matrix form wins by its 2-D register blocking (7 fragment loads per 12 MFMAs).

#include "lsb_internal.h"

namespace lsb {

// ZM_PRESS: out = -acc - (kx^2 + ky^2) aux; ZM_AUXSUB: out = aux - acc (aux a field at the
// output positions, passed in the bct slot; out may alias aux)
enum { ZM_PLAIN = 0, ZM_HELM1 = 1, ZM_HELM2 = 2, ZM_PRESS = 3, ZM_AUXSUB = 4 };

typedef double v4d __attribute__((ext_vector_type(4)));

struct ZArgs {
    const double* M;    // column-major, leading dim ldm; zero rows to ldm, zero columns to round_up(nin, 4)
    int ldm, nout, nin;
    int in_row0;        // first input plane
    int out_row0;       // first output plane
    long P;             // columns per plane (nx*ny)
    int nx, ny;
    int gblocks;        // row blocks per blockIdx.y group
    int xrow;           // a last row past the 16-row blocks (nout % 16 == 1), done on the VALU; -1 if none
    double sigma;
    const double* wavx;   // nx
    const double* wavy;   // ny
    const double* lam;    // nzm (stage 1)
    const double* lift;   // hbot, htop, d2hbot, d2htop (4*nzp)
    const double* elift;  // Einv applied to interior hbot, htop, d2hbot, d2htop (4*nzm)
    int nzp;
};

constexpr int kZRB = 3;   // row blocks (of 16) per wave
constexpr int kZCB = 4;   // column blocks (of 16) per block
constexpr int kZPD = 2;   // k-steps of fragments in flight
constexpr int kZMaxWaves = 8;

// three waves per SIMD (<= 168 VGPRs) keep enough loads in flight to feed the matrix cores
template <int MODE>
__global__ void __launch_bounds__(64 * kZMaxWaves) __attribute__((amdgpu_waves_per_eu(3)))
k_zcontract(ZArgs a, FieldList<double> in, FieldList<double> out, FieldList<double> bct, FieldList<double> bcb) {
    constexpr int RB = kZRB, CB = kZCB, PD = kZPD;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int f = blockIdx.z;
    const int nrb_all = (a.xrow >= 0) ? (a.nout >> 4) : ((a.nout + 15) >> 4);
    const int g0 = blockIdx.y * a.gblocks;
    const int nrb = min(a.gblocks, nrb_all - g0);
    const int base = nrb / nw, extra = nrb % nw;
    const int myn = base + (wave < extra ? 1 : 0);
    if (myn <= 0) return;
    const int rb0 = g0 + wave * base + min(wave, extra);
    const long P = a.P;
    const long p0 = (long)blockIdx.x * (16 * CB);
    long cj[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) cj[j] = min(p0 + 16 * j + li, P - 1);
    const double* __restrict__ X = in.p[f] + (long)a.in_row0 * P;
    const double* __restrict__ Ma = a.M + (long)(rb0 * 16 + li) + (long)lk * a.ldm;
    const int nin = a.nin;
    const int nks = (nin + 3) >> 2;
    v4d acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    // the extra row: wave 0 of the first group accumulates M[xrow][k] X[k][col] on the
    // VALU from the B fragments it holds (lane quarter lk owns k = 4 ks + lk)
    const bool dox = (a.xrow >= 0) && (g0 == 0) && (wave == 0);
    double xacc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) xacc[j] = 0.0;
    double fa[PD][RB], fb[PD][CB];
#pragma unroll
    for (int s = 0; s < PD; ++s) {
        const int ks = min(s, nks - 1);
        const long kr = min(4 * ks + lk, nin - 1);
#pragma unroll
        for (int i = 0; i < RB; ++i) fa[s][i] = Ma[(long)(4 * ks) * a.ldm + 16 * i];
#pragma unroll
        for (int j = 0; j < CB; ++j) fb[s][j] = X[kr * P + cj[j]];
    }
    for (int k0 = 0; k0 < nks; k0 += PD) {
#pragma unroll
        for (int s = 0; s < PD; ++s) {
            const int ks = k0 + s;
            if (ks < nks) {
                double ca[RB], cb[CB];
#pragma unroll
                for (int i = 0; i < RB; ++i) ca[i] = fa[s][i];
#pragma unroll
                for (int j = 0; j < CB; ++j) cb[j] = fb[s][j];
                const int kn = min(ks + PD, nks - 1);
                const long kr = min(4 * kn + lk, nin - 1);
#pragma unroll
                for (int i = 0; i < RB; ++i) fa[s][i] = Ma[(long)(4 * kn) * a.ldm + 16 * i];
#pragma unroll
                for (int j = 0; j < CB; ++j) fb[s][j] = X[kr * P + cj[j]];
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    if (i < myn)
#pragma unroll
                        for (int j = 0; j < CB; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[i], cb[j], acc[i][j], 0, 0, 0);
                if (dox) {
                    // M[xrow][4 ks + lk]: four wave-uniform loads, one selected per lane quarter
                    const double* mc = a.M + a.xrow + (long)(4 * ks) * a.ldm;
                    const double m0 = mc[0], m1 = mc[a.ldm], m2 = mc[2 * a.ldm], m3 = mc[3 * a.ldm];
                    const double cx = lk == 0 ? m0 : (lk == 1 ? m1 : (lk == 2 ? m2 : m3));
#pragma unroll
                    for (int j = 0; j < CB; ++j) xacc[j] = fma(cx, cb[j], xacc[j]);
                }
            }
        }
    }
    // per-pencil epilogue data
    double dfac[CB], vbt[CB], vbb[CB];
    const bool nhom = (MODE == ZM_HELM1 || MODE == ZM_HELM2) && (bct.p[f] != nullptr);
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        dfac[j] = 0.0;
        vbt[j] = 0.0;
        vbb[j] = 0.0;
        if (MODE != ZM_PLAIN) {
            const int c = (int)cj[j];
            const double wx = a.wavx[c % a.nx], wy = a.wavy[c / a.nx];
            dfac[j] = -(wx * wx) - wy * wy - a.sigma;
            if (nhom) {
                vbt[j] = bct.p[f][c];
                vbb[j] = bcb.p[f][c];
            }
        }
    }
    constexpr bool kAux = (MODE == ZM_PRESS || MODE == ZM_AUXSUB);
    const double* __restrict__ auxf = kAux ? bct.p[f] : nullptr;
    double* __restrict__ dst = out.p[f] + (long)a.out_row0 * P;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
            // every operand of the row block is loaded before its first store: the
            // compiler cannot prove the output disjoint from them, and a load placed
            // after a store would wait for its own round trip
            double ep[4][5], ax[4][CB];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rc = min((rb0 + i) * 16 + lk + 4 * v, a.nout - 1);
                if (MODE == ZM_HELM1) {
                    ep[v][0] = a.lam[rc];
                    if (nhom) {
                        const double* e = a.elift;
                        const int n = a.nout;
                        ep[v][1] = e[rc];
                        ep[v][2] = e[n + rc];
                        ep[v][3] = e[2 * n + rc];
                        ep[v][4] = e[3 * n + rc];
                    }
                }
                if (MODE == ZM_HELM2 && nhom) {
                    ep[v][0] = a.lift[rc + 1];
                    ep[v][1] = a.lift[a.nzp + rc + 1];
                }
                if (kAux)
#pragma unroll
                    for (int j = 0; j < CB; ++j) ax[v][j] = auxf[(long)rc * P + cj[j]];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = (rb0 + i) * 16 + lk + 4 * v;
                if (row < a.nout) {
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        const long col = p0 + 16 * j + li;
                        double x = acc[i][j][v];
                        if (MODE == ZM_HELM1) {
                            if (nhom) {
                                // Einv (q - L(h_bot) bc_bot - L(h_top) bc_top), L(h) = h'' + dfac h,
                                // with Einv h and Einv h'' precomputed per interior row
                                x = x - vbb[j] * (ep[v][3] + dfac[j] * ep[v][1]);
                                x = x - vbt[j] * (ep[v][4] + dfac[j] * ep[v][2]);
                            }
                            x = x / (ep[v][0] + dfac[j]);
                        }
                        if (MODE == ZM_HELM2 && nhom) x = x + ep[v][0] * vbb[j] + ep[v][1] * vbt[j];
                        if (MODE == ZM_PRESS) x = -x + dfac[j] * ax[v][j];
                        if (MODE == ZM_AUXSUB) x = ax[v][j] - x;
                        if (col < P) dst[(long)row * P + col] = x;
                    }
                }
            }
        }
    }
    if (dox && (MODE == ZM_PLAIN || kAux)) {
        double ax[CB];
#pragma unroll
        for (int j = 0; j < CB; ++j) ax[j] = kAux ? auxf[(long)a.xrow * P + cj[j]] : 0.0;
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            double x = xacc[j];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            const long col = p0 + 16 * j + li;
            if (MODE == ZM_PRESS) x = -x + dfac[j] * ax[j];
            if (MODE == ZM_AUXSUB) x = ax[j] - x;
            if (lk == 0 && col < P) dst[(long)a.xrow * P + col] = x;
        }
    }
    if (MODE == ZM_HELM2 && g0 == 0 && wave == 0 && lk == 0) {
        // boundary planes 0 and nz of the full field (out_row0 == 1)
        double* full = out.p[f];
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const long col = p0 + 16 * j + li;
            double v0 = 0.0, vn = 0.0;
            if (nhom) {
                v0 = 0.0 + a.lift[0] * vbb[j] + a.lift[a.nzp + 0] * vbt[j];
                vn = 0.0 + a.lift[a.nzp - 1] * vbb[j] + a.lift[a.nzp + a.nzp - 1] * vbt[j];
            }
            if (col < P) {
                full[col] = v0;
                full[(long)(a.nzp - 1) * P + col] = vn;
            }
        }
    }
}

template <int MODE>
static int launch_z(const ZArgs& a0, const FieldList<double>& in, const FieldList<double>& out,
                    const FieldList<double>& bt, const FieldList<double>& bb, int nf, hipStream_t s) {
    ZArgs a = a0;
    a.xrow = ((MODE == ZM_PLAIN || MODE == ZM_PRESS || MODE == ZM_AUXSUB) && a.nout > 16 && a.nout % 16 == 1)
                 ? a.nout - 1
                 : -1;
    const int nrb = (a.xrow >= 0) ? a.nout / 16 : (a.nout + 15) / 16;
    a.gblocks = std::min(nrb, kZRB * kZMaxWaves);
    const int gy = (nrb + a.gblocks - 1) / a.gblocks;
    const int nw = (a.gblocks + kZRB - 1) / kZRB;
    dim3 grid((unsigned)((a.P + 16 * kZCB - 1) / (16 * kZCB)), (unsigned)gy, (unsigned)nf);
    hipLaunchKernelGGL((k_zcontract<MODE>), grid, dim3(64 * nw), 0, s, a, in, out, bt, bb);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// ------------------------------------------------------------------ FP32
// Mixed form: the same contraction on v_mfma_f32_16x16x4f32 (FP32 products, FP32
// accumulation; twice the FP64 rate).  Matrix and panel fragments are loaded in
// their stored precision and rounded to FP32 in registers; the output is stored
// as TO.  Output lane layout of the FP32 instruction: lane l, element v holds
// D[Z32_ROW(l/16, v)][l%16].
#define Z32_ROW(lk, v) (4 * (lk) + (v))
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kZRB32 = 4;

template <class TI, class TO, bool AUX>
__global__ void __launch_bounds__(64 * kZMaxWaves) k_zcontract_f32(ZArgs a, FieldList<TI> in, FieldList<TO> out,
                                                                    FieldList<TO> aux) {
    constexpr int RB = kZRB32, CB = kZCB, PD = kZPD;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int f = blockIdx.z;
    const int nrb_all = (a.xrow >= 0) ? (a.nout >> 4) : ((a.nout + 15) >> 4);
    const int g0 = blockIdx.y * a.gblocks;
    const int nrb = min(a.gblocks, nrb_all - g0);
    const int base = nrb / nw, extra = nrb % nw;
    const int myn = base + (wave < extra ? 1 : 0);
    if (myn <= 0) return;
    const int rb0 = g0 + wave * base + min(wave, extra);
    const long P = a.P;
    const long p0 = (long)blockIdx.x * (16 * CB);
    long cj[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) cj[j] = min(p0 + 16 * j + li, P - 1);
    const TI* __restrict__ X = in.p[f] + (long)a.in_row0 * P;
    const double* __restrict__ Ma = a.M + (long)(rb0 * 16 + li) + (long)lk * a.ldm;
    const int nin = a.nin;
    const int nks = (nin + 3) >> 2;
    v4f acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
    const bool dox = (a.xrow >= 0) && (g0 == 0) && (wave == 0);
    const double* __restrict__ Mx = a.M + (dox ? a.xrow : 0) + (long)lk * a.ldm;
    float xacc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) xacc[j] = 0.f;
    float fa[PD][RB], fb[PD][CB], fx[PD];
#pragma unroll
    for (int s = 0; s < PD; ++s) {
        const int ks = min(s, nks - 1);
        const long kr = min(4 * ks + lk, nin - 1);
#pragma unroll
        for (int i = 0; i < RB; ++i) fa[s][i] = (float)Ma[(long)(4 * ks) * a.ldm + 16 * i];
#pragma unroll
        for (int j = 0; j < CB; ++j) fb[s][j] = (float)X[kr * P + cj[j]];
        fx[s] = dox ? (float)Mx[(long)(4 * ks) * a.ldm] : 0.f;
    }
    for (int k0 = 0; k0 < nks; k0 += PD) {
#pragma unroll
        for (int s = 0; s < PD; ++s) {
            const int ks = k0 + s;
            if (ks < nks) {
                float ca[RB], cb[CB];
#pragma unroll
                for (int i = 0; i < RB; ++i) ca[i] = fa[s][i];
#pragma unroll
                for (int j = 0; j < CB; ++j) cb[j] = fb[s][j];
                const float cx = fx[s];
                const int kn = min(ks + PD, nks - 1);
                const long kr = min(4 * kn + lk, nin - 1);
#pragma unroll
                for (int i = 0; i < RB; ++i) fa[s][i] = (float)Ma[(long)(4 * kn) * a.ldm + 16 * i];
#pragma unroll
                for (int j = 0; j < CB; ++j) fb[s][j] = (float)X[kr * P + cj[j]];
                fx[s] = dox ? (float)Mx[(long)(4 * kn) * a.ldm] : 0.f;
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    if (i < myn)
#pragma unroll
                        for (int j = 0; j < CB; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[i], cb[j], acc[i][j], 0, 0, 0);
                if (dox)
#pragma unroll
                    for (int j = 0; j < CB; ++j) xacc[j] = fmaf(cx, cb[j], xacc[j]);
            }
        }
    }
    const TO* __restrict__ auxf = AUX ? aux.p[f] : nullptr;
    TO* __restrict__ dst = out.p[f] + (long)a.out_row0 * P;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
            // the row block's aux operands are loaded before its first store (see the
            // FP64 kernel)
            float ax[4][CB];
            if (AUX)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rc = min((rb0 + i) * 16 + Z32_ROW(lk, v), a.nout - 1);
#pragma unroll
                    for (int j = 0; j < CB; ++j) ax[v][j] = (float)auxf[(long)rc * P + cj[j]];
                }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = (rb0 + i) * 16 + Z32_ROW(lk, v);
                if (row < a.nout) {
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        const long col = p0 + 16 * j + li;
                        float x = acc[i][j][v];
                        if (AUX) x = ax[v][j] - x;
                        if (col < P) dst[(long)row * P + col] = (TO)x;
                    }
                }
            }
        }
    }
    if (dox) {
        float ax[CB];
#pragma unroll
        for (int j = 0; j < CB; ++j) ax[j] = AUX ? (float)auxf[(long)a.xrow * P + cj[j]] : 0.f;
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            float x = xacc[j];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            const long col = p0 + 16 * j + li;
            if (AUX) x = ax[j] - x;
            if (lk == 0 && col < P) dst[(long)a.xrow * P + col] = (TO)x;
        }
    }
}

template <class TI, class TO>
int zcontract_f32(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const TI* const* in, TO* const* out,
                  int nf, hipStream_t s, long P, const TO* const* aux) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("zcontract: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    ZArgs a{};
    a.M = M;
    a.ldm = ldm;
    a.nout = nout;
    a.nin = nin;
    a.P = P > 0 ? P : (long)c->plane_s;
    FieldList<TI> fi{};
    FieldList<TO> fo{}, fa{};
    for (int f = 0; f < nf; ++f) {
        fi.p[f] = const_cast<TI*>(in[f]);
        fo.p[f] = out[f];
        fa.p[f] = aux ? const_cast<TO*>(aux[f]) : nullptr;
    }
    a.xrow = (nout > 16 && nout % 16 == 1) ? nout - 1 : -1;
    const int nrb = (a.xrow >= 0) ? nout / 16 : (nout + 15) / 16;
    a.gblocks = std::min(nrb, kZRB32 * kZMaxWaves);
    const int gy = (nrb + a.gblocks - 1) / a.gblocks;
    const int nw = (a.gblocks + kZRB32 - 1) / kZRB32;
    dim3 grid((unsigned)((a.P + 16 * kZCB - 1) / (16 * kZCB)), (unsigned)gy, (unsigned)nf);
    int tok = -1;
    prof_mark_begin(c, PC_Z, s, &tok, 2.0 * nout * nin * (double)a.P * nf,
                    (double)a.P * nf * (nin * sizeof(TI) + nout * sizeof(TO) * (aux ? 2 : 1)));
    if (aux)
        hipLaunchKernelGGL((k_zcontract_f32<TI, TO, true>), grid, dim3(64 * nw), 0, s, a, fi, fo, fa);
    else
        hipLaunchKernelGGL((k_zcontract_f32<TI, TO, false>), grid, dim3(64 * nw), 0, s, a, fi, fo, fa);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}
template int zcontract_f32<double, float>(lsb_ctx*, const double*, int, int, int, const double* const*,
                                          float* const*, int, hipStream_t, long, const float* const*);
template int zcontract_f32<double, double>(lsb_ctx*, const double*, int, int, int, const double* const*,
                                           double* const*, int, hipStream_t, long, const double* const*);
template int zcontract_f32<float, float>(lsb_ctx*, const double*, int, int, int, const float* const*,
                                         float* const*, int, hipStream_t, long, const float* const*);

int zcontract(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const double* const* in,
              double* const* out, int nf, hipStream_t s) {
    return zcontract_ex(c, M, ldm, nout, nin, in, out, nf, (long)c->plane_s, nullptr, s);
}

// P columns per plane (spectral nx*ny or physical (nx+2)*ny); with aux, out = aux - M in
int zcontract_ex(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const double* const* in,
                 double* const* out, int nf, long P, const double* const* aux, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("zcontract: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    ZArgs a{};
    a.M = M;
    a.ldm = ldm;
    a.nout = nout;
    a.nin = nin;
    a.P = P;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nzp = c->nzp;
    FieldList<double> fi{}, fo{}, fa{}, z{};
    for (int f = 0; f < nf; ++f) {
        fi.p[f] = const_cast<double*>(in[f]);
        fo.p[f] = out[f];
        fa.p[f] = aux ? const_cast<double*>(aux[f]) : nullptr;
        z.p[f] = nullptr;
    }
    int tok = -1;
    prof_mark_begin(c, PC_Z, s, &tok, 2.0 * nout * nin * (double)a.P * nf,
                    (double)(nin + nout + (aux ? nout : 0)) * a.P * nf * sizeof(double));
    int rc = aux ? launch_z<ZM_AUXSUB>(a, fi, fo, fa, z, nf, s) : launch_z<ZM_PLAIN>(a, fi, fo, z, z, nf, s);
    prof_mark_end(c, tok, s);
    return rc;
}

// Pressure right-hand side (navierstokes.f90:1411-1418) from div_h = ikx u + iky v:
// rhs = -d(div_h)/dz - (kx^2 + ky^2) w.  d/dz commutes with the horizontal
// divergence, so this is the reference's divuv(dfdz(u), dfdz(v)) with one
// contraction instead of two.
int zcontract_press(lsb_ctx* c, const double* divh, const double* w, double* rhs, hipStream_t s) {
    ZArgs a{};
    a.M = c->d_mdz;
    a.ldm = c->ldm_z;
    a.nout = c->nzp;
    a.nin = c->nzp;
    a.P = (long)c->plane_s;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nzp = c->nzp;
    a.sigma = 0.0;
    a.wavx = c->d_wavx;
    a.wavy = c->d_wavy;
    FieldList<double> fi{}, fo{}, aux{}, z{};
    fi.p[0] = const_cast<double*>(divh);
    fo.p[0] = rhs;
    aux.p[0] = const_cast<double*>(w);
    z.p[0] = nullptr;
    int tok = -1;
    prof_mark_begin(c, PC_Z, s, &tok, 2.0 * c->nzp * (double)c->nzp * a.P, 3.0 * c->nzp * (double)a.P * sizeof(double));
    int rc = launch_z<ZM_PRESS>(a, fi, fo, aux, z, 1, s);
    prof_mark_end(c, tok, s);
    return rc;
}

// In-place Helmholtz/Poisson solves of nf fields (interior rows 1..nz-1), with
// optional non-homogeneous Dirichlet data bc_top (plane 0) / bc_bot (plane nzp-1).
int helmholtz(lsb_ctx* c, double* const* rhs, const double* const* bc_top, const double* const* bc_bot, int nf,
              double sigma, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("helmholtz: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    const int nzm = c->nzm;
    const size_t tbytes = (size_t)nf * nzm * c->plane_s * sizeof(double);
    void* tbuf = nullptr;
    LSB_TRY(stage_get(c, 7, tbytes, &tbuf));
    FieldList<double> fr{}, ft{}, bt{}, bb{};
    for (int f = 0; f < nf; ++f) {
        fr.p[f] = rhs[f];
        ft.p[f] = (double*)tbuf + (size_t)f * nzm * c->plane_s;
        bt.p[f] = bc_top ? const_cast<double*>(bc_top[f]) : nullptr;
        bb.p[f] = bc_bot ? const_cast<double*>(bc_bot[f]) : nullptr;
    }
    ZArgs a{};
    a.ldm = ((nzm + 63) / 64) * 64 + 64;
    a.nout = nzm;
    a.nin = nzm;
    a.P = (long)c->plane_s;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nzp = c->nzp;
    a.sigma = sigma;
    a.wavx = c->d_wavx;
    a.wavy = c->d_wavy;
    a.lam = c->d_diag;
    a.lift = c->d_lift;
    a.elift = c->d_lift + 4 * c->nzp;
    int tok = -1;
    // two nzm x nzm projections per pencil; read rhs and write the solution once each
    prof_mark_begin(c, PC_HELM, s, &tok, 4.0 * nzm * (double)nzm * a.P * nf,
                    2.0 * c->nzp * (double)a.P * nf * sizeof(double));
    a.M = c->d_simat;
    a.in_row0 = 1;
    a.out_row0 = 0;
    int rc = launch_z<ZM_HELM1>(a, fr, ft, bt, bb, nf, s);
    if (rc == LSB_OK) {
        a.M = c->d_smat;
        a.in_row0 = 0;
        a.out_row0 = 1;
        rc = launch_z<ZM_HELM2>(a, ft, fr, bt, bb, nf, s);
    }
    prof_mark_end(c, tok, s);
    return rc;
}

}  // namespace lsb
