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

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace lsb {

// ZM_PRESS: out = -acc - (kx^2 + ky^2) aux; ZM_AUXSUB: out = aux - acc (aux a field at the
// output positions, passed in the bct slot; out may alias aux)
// Parity-split kernel only:
// ZM_PRESSDIV: out = -M div_h - (kx^2 + ky^2) aux with div_h = i kx u + i ky v formed as
//   the panel is loaded (in.p[0] = u, in.p[1] = v, aux = w);
// ZM_PROJ: the projection of navierstokes.f90:1466-1488 applied to y = M phi as it is
//   stored: p = -(y + div_h)/(k^2 dt bdfv), u -= dt bdfv i kx p, v -= dt bdfv i ky p
//   (out.p[0] = p, bct.p[0] = u, bct.p[1] = v, updated in place).
enum { ZM_PLAIN = 0, ZM_HELM1 = 1, ZM_HELM2 = 2, ZM_PRESS = 3, ZM_AUXSUB = 4, ZM_PRESSDIV = 5, ZM_PROJ = 6 };

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
    // parity-split contractions: M holds P, M2 holds Q; nsp rows in the physical
    // (or parity-ordered mode) layout; osgn the reflection sign of the output
    const double* M2;
    int nsp;
    double osgn;
    const float* M32;   // FP32 copy of [P | Q] for the FP32 split kernel, Q at M32 + q32
    long q32;
    double pdt, pbdfv;  // ZM_PROJ: dt and bdfv
    const double* M3;   // k_zhelm: stage-2 P and Q (stage 1 in M, M2)
    const double* M4;
    const float* M32b;  // k_zhelm32: FP32 stage-2 P, Q at M32b + q32b (stage 1 in M32, q32)
    long q32b;
    const double* eel;  // k_zhelm32: E applied to the four Einv-lifting vectors (4*nzm, physical rows)
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

// ------------------------------------------------------------ parity split
// On a grid symmetric about the mid-plane (Gauss-Lobatto points under mapping 0
// or 3) the z operators are centro-symmetric or centro-antisymmetric
// (M[n-1-i][n-1-j] = sgn M[i][j]: d2 and E, Einv by modes; d with sgn = -1) and
// the Helmholtz eigenvectors are each even or odd.  With s_j = x_j + x_{n-1-j},
// d_j = x_j - x_{n-1-j} (j < me = (n+1)/2; the middle point of odd n gives
// s = 2 x_m, d = 0):
//   E_i = sum_j P[i][j] s_j,  O_i = sum_j Q[i][j] d_j,
//   P[i][j] = (M[i][j] + M[i][n-1-j]) / 2,  Q[i][j] = (M[i][j] - M[i][n-1-j]) / 2,
//   y_i = E_i + O_i,  y_{n-1-i} = sgn (E_i - O_i)          (i < me)
// -- two half-size products, half the matrix-core work of the dense one.  The
// Helmholtz mode space is kept in parity order (even mode a in row a, odd mode b
// in row n-1-b): stage 1 (Einv) folds its input and writes E_i, O_i to those rows
// unmixed; stage 2 (E) reads them unmixed and unfolds into physical rows.  A wave
// owns the same rows of both halves (two accumulator sets, two column blocks).
constexpr int kZSCB = 2;
// FP32 split: 3 x 4 tiles per set at two waves per SIMD (measured best of 3-4 row
// by 1-8 column blocks on the 193-row physical-space contraction)
constexpr int kZSRB32 = 3, kZSCB32 = 4;
constexpr int kZSRB = 3;  // FP64 row blocks per wave (two 3 x 2 tile sets, 96 accumulator VGPRs)

// element of a field or matrix at a 32-bit element offset from a wave-uniform base
// (the address is base + zero-extended byte offset: the load takes the base in
// scalar registers)
template <class T>
__device__ __forceinline__ T ldo(const T* __restrict__ base, unsigned eoff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(eoff * (unsigned)sizeof(T)));
}

template <int MODE, bool XR, int RB = kZSRB, int CB = kZSCB, int PD = kZPD, int WPE = 3>
__global__ void __launch_bounds__(64 * kZMaxWaves) __attribute__((amdgpu_waves_per_eu(WPE)))
k_zsplit(ZArgs a, FieldList<double> in, FieldList<double> out, FieldList<double> bct, FieldList<double> bcb) {
    constexpr bool kInMix = (MODE != ZM_HELM2);   // fold a physical input
    constexpr bool kOutMix = (MODE != ZM_HELM1);  // unfold into physical rows
    constexpr bool kAux = (MODE == ZM_PRESS || MODE == ZM_AUXSUB || MODE == ZM_PRESSDIV);
    constexpr bool kDiv = (MODE == ZM_PRESSDIV), kProj = (MODE == ZM_PROJ);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int f = blockIdx.z;
    const int me = a.nout, n = a.nsp, mo = n - me;
    const int nrb_all = XR ? (me >> 4) : ((me + 15) >> 4);
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
    // The k loop has no branches, so that the loads of k-step s + PD stay in flight
    // while step s multiplies: the k-steps are padded to a multiple of PD (the matrix
    // has zero columns up to a multiple of 48), all RB row blocks are multiplied (rows
    // past the wave's share are zero or unused), the prefetch position is clamped
    // (matrix: last padded step; panel: last k-step whose 4 rows fit), and the middle
    // row (XR) is accumulated by every wave from per-lane loads.  Offsets are 32-bit
    // element offsets (a field holds < 2^29 elements).
    const double* __restrict__ X = in.p[f] + (long)a.in_row0 * P;
    const double* __restrict__ Xv = kDiv ? in.p[1] + (long)a.in_row0 * P : nullptr;
    const double* __restrict__ Mp = a.M;
    const double* __restrict__ Mq = a.M2;
    const unsigned xstep = 4u * (unsigned)P, mstep = 4u * (unsigned)a.ldm;
    // the pencil's wavenumbers; sgc = -1 on a real column, +1 on an imaginary one, so
    // that (i k x)_c = sgc k x_(c^1) (derivatives.f90:306-326 divuv)
    double wxc[CB], wyc[CB], sgc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        wxc[j] = 0.0;
        wyc[j] = 0.0;
        sgc[j] = 0.0;
        if (kDiv || kProj) {
            const int c = (int)cj[j];
            wxc[j] = a.wavx[2 * ((c % a.nx) >> 1)];
            wyc[j] = a.wavy[c / a.nx];
            sgc[j] = (c & 1) ? 1.0 : -1.0;
        }
    }
    unsigned ot[CB], ob[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const unsigned cl = (unsigned)(kDiv ? (cj[j] ^ 1) : cj[j]);  // div_h reads the partner column
        ot[j] = (unsigned)lk * (unsigned)P + cl;
        ob[j] = (unsigned)(n - 1 - lk) * (unsigned)P + cl;
    }
    const unsigned om = (unsigned)(rb0 * 16 + li + lk * a.ldm);
    const unsigned omx = (unsigned)((XR ? a.xrow : 0) + lk * a.ldm);
    const int nks = (me + 3) >> 2;
    const int nkp = (nks + PD - 1) / PD * PD;
    const int kmax = (n - 4) >> 2;
    v4d accE[RB][CB], accO[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            accE[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
            accO[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
        }
    double xacc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) xacc[j] = 0.0;
    double fa[PD][RB], fq[PD][RB], fx[PD][CB], fy[PD][CB], gx[PD][CB], gy[PD][CB], fxp[PD], fxq[PD];
    auto load = [&](int s, int k) {
        const unsigned mk = (unsigned)min(k, nkp - 1) * mstep;
        const unsigned xk = (unsigned)min(k, kmax) * xstep;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            fa[s][i] = ldo(Mp, om + mk + 16u * i);
            fq[s][i] = ldo(Mq, om + mk + 16u * i);
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            fx[s][j] = ldo(X, ot[j] + xk);
            fy[s][j] = ldo(X, ob[j] - xk);
            if (kDiv) {
                gx[s][j] = ldo(Xv, ot[j] + xk);
                gy[s][j] = ldo(Xv, ob[j] - xk);
            }
        }
        if (XR) {
            fxp[s] = ldo(Mp, omx + mk);
            fxq[s] = ldo(Mq, omx + mk);
        }
        // keeps the prefetch where it is written: without the asm the compiler turns
        // the phi of the two loads (prologue / previous step) into one load at the
        // use; without the scheduling barrier the loads sink below the next step's
        // products
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
    };
#pragma unroll
    for (int s = 0; s < PD; ++s) load(s, s);
    for (int k0 = 0; k0 < nkp; k0 += PD) {
#pragma unroll
        for (int s = 0; s < PD; ++s) {
            double sv[CB], dv[CB];
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                double xt = fx[s][j], xb = fy[s][j];
                if (kDiv) {
                    xt = sgc[j] * (wxc[j] * fx[s][j] + wyc[j] * gx[s][j]);
                    xb = sgc[j] * (wxc[j] * fy[s][j] + wyc[j] * gy[s][j]);
                }
                sv[j] = kInMix ? xt + xb : xt;
                dv[j] = kInMix ? xt - xb : xb;
            }
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    accE[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s][i], sv[j], accE[i][j], 0, 0, 0);
                    accO[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fq[s][i], dv[j], accO[i][j], 0, 0, 0);
                }
            if (XR)
#pragma unroll
                for (int j = 0; j < CB; ++j) xacc[j] = fma(fxq[s], dv[j], fma(fxp[s], sv[j], xacc[j]));
            load(s, k0 + s + PD);
        }
    }
    const bool dox = XR && (g0 == 0) && (wave == 0);
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
    const double* __restrict__ auxf = kAux ? bct.p[f] : nullptr;
    double* __restrict__ dst = out.p[f] + (long)a.out_row0 * P;
    const double osgn = a.osgn;
    double* __restrict__ pu = kProj ? bct.p[0] : nullptr;
    double* __restrict__ pv = kProj ? bct.p[1] : nullptr;
    // ZM_PROJ store of y at (row, col): every lane of a 16-lane quarter calls it (the
    // partner column c^1 is the neighbouring lane), stores are masked by `st`
    auto proj_store = [&](long row, int j, double y, double uc, double vc, bool st) {
        const double up = __shfl_xor(uc, 1), vp = __shfl_xor(vc, 1);
        const double divf = sgc[j] * (wxc[j] * up + wyc[j] * vp);
        const double wav2 = wxc[j] * wxc[j] + wyc[j] * wyc[j];
        double pc = 0.0;
        if (wav2 != 0.0) pc = -(y + divf) / (wav2 * a.pdt * a.pbdfv);
        const double pp = __shfl_xor(pc, 1);
        const double dtb = a.pdt * a.pbdfv;
        if (st) {
            const long o = row * P + cj[j];
            dst[o] = pc;
            pu[o] = uc - dtb * (sgc[j] * wxc[j] * pp);
            pv[o] = vc - dtb * (sgc[j] * wyc[j] * pp);
        }
    };
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                // the stored row of (row block i, element v) in this half: r or n-1-r
                // (operands are gathered before the first store, as in k_zcontract)
                double ep[4][5], ax[4][CB], ay[4][CB];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rb0 + i) * 16 + lk + 4 * v, me - 1);
                    const int rc = half ? n - 1 - r : r;
                    if (kProj)
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            ax[v][j] = pu[(long)rc * P + cj[j]];
                            ay[v][j] = pv[(long)rc * P + cj[j]];
                        }
                    if (MODE == ZM_HELM1) {
                        ep[v][0] = a.lam[rc];
                        if (nhom) {
                            ep[v][1] = a.elift[rc];
                            ep[v][2] = a.elift[n + rc];
                            ep[v][3] = a.elift[2 * n + rc];
                            ep[v][4] = a.elift[3 * n + rc];
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
                    const int r = (rb0 + i) * 16 + lk + 4 * v;
                    // the bottom half holds rows r < mo only (the middle row of odd n,
                    // and the even-only last mode, have no mirror)
                    if (r < me && (half == 0 || r < mo)) {
                        const int row = half ? n - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            const long col = p0 + 16 * j + li;
                            const double e = accE[i][j][v], o = accO[i][j][v];
                            double x;
                            if (kOutMix)
                                x = half ? osgn * (e - o) : e + o;
                            else
                                x = half ? o : e;
                            if (MODE == ZM_HELM1) {
                                if (nhom) {
                                    x = x - vbb[j] * (ep[v][3] + dfac[j] * ep[v][1]);
                                    x = x - vbt[j] * (ep[v][4] + dfac[j] * ep[v][2]);
                                }
                                x = x / (ep[v][0] + dfac[j]);
                            }
                            if (MODE == ZM_HELM2 && nhom) x = x + ep[v][0] * vbb[j] + ep[v][1] * vbt[j];
                            if (MODE == ZM_PRESS || kDiv) x = -x + dfac[j] * ax[v][j];
                            if (MODE == ZM_AUXSUB) x = ax[v][j] - x;
                            if (kProj)
                                proj_store(row, j, x, ax[v][j], ay[v][j], col < P);
                            else if (col < P)
                                dst[(long)row * P + col] = x;
                        }
                    }
                }
            }
        }
    }
    if (dox && (MODE == ZM_PLAIN || kAux || kProj)) {
        double ax[CB], ay[CB];
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            ax[j] = kAux ? auxf[(long)a.xrow * P + cj[j]] : (kProj ? pu[(long)a.xrow * P + cj[j]] : 0.0);
            ay[j] = kProj ? pv[(long)a.xrow * P + cj[j]] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            double x = xacc[j];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            const long col = p0 + 16 * j + li;
            if (MODE == ZM_PRESS || kDiv) x = -x + dfac[j] * ax[j];
            if (MODE == ZM_AUXSUB) x = ax[j] - x;
            if (kProj)
                proj_store(a.xrow, j, x, ax[j], ay[j], lk == 0 && col < P);
            else if (lk == 0 && col < P)
                dst[(long)a.xrow * P + col] = x;
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

// FP32 form of k_zsplit for the mixed nonlinear chain (folded physical input,
// unfolded output; aux subtraction as in k_zcontract_f32), same branch-free
// pipelined k loop; the matrix halves are read from their FP32 copies.
template <class TI, class TO, bool AUX, bool XR, int RB = kZSRB32, int CB = kZSCB32, int PD = kZPD, int WPE = 2>
__global__ void __launch_bounds__(64 * kZMaxWaves) __attribute__((amdgpu_waves_per_eu(WPE)))
k_zsplit_f32(ZArgs a, FieldList<TI> in, FieldList<TO> out, FieldList<TO> aux) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int f = blockIdx.z;
    const int me = a.nout, n = a.nsp, mo = n - me;
    const int nrb_all = XR ? (me >> 4) : ((me + 15) >> 4);
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
    const float* __restrict__ Mp = a.M32;
    const float* __restrict__ Mq = a.M32 + a.q32;
    const unsigned xstep = 4u * (unsigned)P, mstep = 4u * (unsigned)a.ldm;
    unsigned ot[CB], ob[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        ot[j] = (unsigned)lk * (unsigned)P + (unsigned)cj[j];
        ob[j] = (unsigned)(n - 1 - lk) * (unsigned)P + (unsigned)cj[j];
    }
    const unsigned om = (unsigned)(rb0 * 16 + li + lk * a.ldm);
    const unsigned omx = (unsigned)((XR ? a.xrow : 0) + lk * a.ldm);
    const int nks = (me + 3) >> 2;
    const int nkp = (nks + PD - 1) / PD * PD;
    const int kmax = (n - 4) >> 2;
    v4f accE[RB][CB], accO[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            accE[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
            accO[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    float xacc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) xacc[j] = 0.f;
    float fa[PD][RB], fq[PD][RB], fxp[PD], fxq[PD];
    TI fx[PD][CB], fy[PD][CB];
    auto load = [&](int s, int k) {
        const unsigned mk = (unsigned)min(k, nkp - 1) * mstep;
        const unsigned xk = (unsigned)min(k, kmax) * xstep;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            fa[s][i] = ldo(Mp, om + mk + 16u * i);
            fq[s][i] = ldo(Mq, om + mk + 16u * i);
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            fx[s][j] = ldo(X, ot[j] + xk);
            fy[s][j] = ldo(X, ob[j] - xk);
        }
        if (XR) {
            fxp[s] = ldo(Mp, omx + mk);
            fxq[s] = ldo(Mq, omx + mk);
        }
        asm volatile("" ::: "memory");  // keep the prefetch in place (see k_zsplit)
        __builtin_amdgcn_sched_barrier(0);
    };
#pragma unroll
    for (int s = 0; s < PD; ++s) load(s, s);
    for (int k0 = 0; k0 < nkp; k0 += PD) {
#pragma unroll
        for (int s = 0; s < PD; ++s) {
            float sv[CB], dv[CB];
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const float x0 = (float)fx[s][j], x1 = (float)fy[s][j];
                sv[j] = x0 + x1;
                dv[j] = x0 - x1;
            }
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    accE[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[s][i], sv[j], accE[i][j], 0, 0, 0);
                    accO[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fq[s][i], dv[j], accO[i][j], 0, 0, 0);
                }
            if (XR)
#pragma unroll
                for (int j = 0; j < CB; ++j) xacc[j] = fmaf(fxq[s], dv[j], fmaf(fxp[s], sv[j], xacc[j]));
            load(s, k0 + s + PD);
        }
    }
    const bool dox = XR && (g0 == 0) && (wave == 0);
    const TO* __restrict__ auxf = AUX ? aux.p[f] : nullptr;
    TO* __restrict__ dst = out.p[f] + (long)a.out_row0 * P;
    const float osgn = (float)a.osgn;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                float ax[4][CB];
                if (AUX)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = min((rb0 + i) * 16 + Z32_ROW(lk, v), me - 1);
                        const int rc = half ? n - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) ax[v][j] = (float)auxf[(long)rc * P + cj[j]];
                    }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rb0 + i) * 16 + Z32_ROW(lk, v);
                    if (r < me && (half == 0 || r < mo)) {
                        const int row = half ? n - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            const long col = p0 + 16 * j + li;
                            const float e = accE[i][j][v], o = accO[i][j][v];
                            float x = half ? osgn * (e - o) : e + o;
                            if (AUX) x = ax[v][j] - x;
                            if (col < P) dst[(long)row * P + col] = (TO)x;
                        }
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

// Row-block grouping shared by the split launches; the middle row of odd n goes to
// the VALU when me % 16 == 1 (modes without a per-row epilogue only).
static void zsplit_shape(ZArgs& a, const ZSplitOp& op, bool allow_xrow, int rb, int* gy, int* nw) {
    a.M = op.d_m;
    a.M2 = op.d_m + op.qoff;
    a.M32 = op.d_m32;
    a.q32 = (long)op.qoff;
    a.ldm = op.ldm;
    a.nout = op.me;
    a.nin = op.me;
    a.nsp = op.n;
    a.osgn = op.osgn;
    a.xrow = (allow_xrow && (op.n & 1) && op.me > 16 && op.me % 16 == 1) ? op.me - 1 : -1;
    const int nrb = (a.xrow >= 0) ? op.me / 16 : (op.me + 15) / 16;
    a.gblocks = std::min(nrb, rb * kZMaxWaves);
    *gy = (nrb + a.gblocks - 1) / a.gblocks;
    *nw = (a.gblocks + rb - 1) / rb;
}

template <int MODE>
static int launch_zs(const ZArgs& a0, const ZSplitOp& op, const FieldList<double>& in, const FieldList<double>& out,
                     const FieldList<double>& bt, const FieldList<double>& bb, int nf, hipStream_t s) {
    ZArgs a = a0;
    int gy = 1, nw = 1;
    // two input fields or a wider epilogue: two row blocks per wave keep three waves per SIMD
    constexpr int RB = (MODE == ZM_PRESSDIV || MODE == ZM_PROJ) ? 2 : kZSRB;
    zsplit_shape(a, op, MODE != ZM_HELM1 && MODE != ZM_HELM2, RB, &gy, &nw);
    dim3 grid((unsigned)((a.P + 16 * kZSCB - 1) / (16 * kZSCB)), (unsigned)gy, (unsigned)nf);
    if (a.xrow >= 0)
        hipLaunchKernelGGL((k_zsplit<MODE, true, RB>), grid, dim3(64 * nw), 0, s, a, in, out, bt, bb);
    else
        hipLaunchKernelGGL((k_zsplit<MODE, false, RB>), grid, dim3(64 * nw), 0, s, a, in, out, bt, bb);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// the split kernels address a field with 32-bit element offsets
static bool zs_fits(const ZSplitOp& op, long P) { return op.ok && (long)op.n * P < (1L << 29) && op.n >= 8; }

// Both Helmholtz projections of a block's pencils in one kernel: stage 1 (folded
// rhs, Einv halves, lifting and 1/(lam - k^2 - sigma)) leaves the parity-ordered
// mode values of its 16 CB pencils in LDS; after a barrier stage 2 (E halves, h bc,
// boundary planes) reads them from there.  The mode-space buffer of the two-kernel
// form (two HBM passes over nzm planes per field) disappears.  The block owns its
// pencils for all rows, so the in-place update of rhs is race-free.
template <int RB, int CB, int PD, int WPE>
__global__ void __launch_bounds__(64 * kZMaxWaves) __attribute__((amdgpu_waves_per_eu(WPE)))
k_zhelm(ZArgs a, FieldList<double> fld, FieldList<double> bct, FieldList<double> bcb) {
    extern __shared__ double T[];  // n rows x W pencils
    constexpr int W = 16 * CB;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int f = blockIdx.z;
    const int me = a.nout, n = a.nsp, mo = n - me;
    const int nrb = (me + 15) >> 4;
    const int base = nrb / nw, extra = nrb % nw;
    const int myn = base + (wave < extra ? 1 : 0);
    const int rb0 = wave * base + min(wave, extra);
    const long P = a.P;
    const long p0 = (long)blockIdx.x * W;
    long cj[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) cj[j] = min(p0 + 16 * j + li, P - 1);
    double* __restrict__ Fq = fld.p[f];
    const double* __restrict__ X = Fq + P;  // interior rows 1..nz
    const unsigned xstep = 4u * (unsigned)P, mstep = 4u * (unsigned)a.ldm;
    unsigned ot[CB], ob[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        ot[j] = (unsigned)lk * (unsigned)P + (unsigned)cj[j];
        ob[j] = (unsigned)(n - 1 - lk) * (unsigned)P + (unsigned)cj[j];
    }
    const unsigned om = (unsigned)(rb0 * 16 + li + lk * a.ldm);
    const int nks = (me + 3) >> 2;
    const int nkp = (nks + PD - 1) / PD * PD;
    const int kmax = (n - 4) >> 2;
    // per-pencil data of both stages
    double dfac[CB], vbt[CB], vbb[CB];
    const bool nhom = bct.p[f] != nullptr;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int c = (int)cj[j];
        const double wx = a.wavx[c % a.nx], wy = a.wavy[c / a.nx];
        dfac[j] = -(wx * wx) - wy * wy - a.sigma;
        vbt[j] = nhom ? bct.p[f][c] : 0.0;
        vbb[j] = nhom ? bcb.p[f][c] : 0.0;
    }
    v4d accE[RB][CB], accO[RB][CB];
    double fa[PD][RB], fq[PD][RB], fx[PD][CB], fy[PD][CB];
    // ---------------------------------------------------------------- stage 1
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            accE[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
            accO[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
        }
    {
        const double* __restrict__ Mp = a.M;
        const double* __restrict__ Mq = a.M2;
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkp - 1) * mstep;
            const unsigned xk = (unsigned)min(k, kmax) * xstep;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(Mp, om + mk + 16u * i);
                fq[s][i] = ldo(Mq, om + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = ldo(X, ot[j] + xk);
                fy[s][j] = ldo(X, ob[j] - xk);
            }
            asm volatile("" ::: "memory");  // keep the prefetch in place (see k_zsplit)
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkp; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                double sv[CB], dv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    sv[j] = fx[s][j] + fy[s][j];
                    dv[j] = fx[s][j] - fy[s][j];
                }
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        accE[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s][i], sv[j], accE[i][j], 0, 0, 0);
                        accO[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fq[s][i], dv[j], accO[i][j], 0, 0, 0);
                    }
                load(s, k0 + s + PD);
            }
        }
    }
    // stage-1 epilogue into LDS: even mode r -> row r, odd mode r -> row n-1-r
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double ep[4][5];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rb0 + i) * 16 + lk + 4 * v, me - 1);
                    const int rc = half ? n - 1 - r : r;
                    ep[v][0] = a.lam[rc];
                    if (nhom) {
                        ep[v][1] = a.elift[rc];
                        ep[v][2] = a.elift[n + rc];
                        ep[v][3] = a.elift[2 * n + rc];
                        ep[v][4] = a.elift[3 * n + rc];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rb0 + i) * 16 + lk + 4 * v;
                    if (r < me && (half == 0 || r < mo)) {
                        const int row = half ? n - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            double x = half ? accO[i][j][v] : accE[i][j][v];
                            if (nhom) {
                                x = x - vbb[j] * (ep[v][3] + dfac[j] * ep[v][1]);
                                x = x - vbt[j] * (ep[v][4] + dfac[j] * ep[v][2]);
                            }
                            x = x / (ep[v][0] + dfac[j]);
                            T[row * W + 16 * j + li] = x;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // ---------------------------------------------------------------- stage 2
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            accE[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
            accO[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
        }
    {
        const double* __restrict__ Mp = a.M3;
        const double* __restrict__ Mq = a.M4;
        // mode values from LDS: even mode kr in row kr, odd mode kr in row n-1-kr (rows
        // past me - 1 meet zero matrix columns; clamped to stay in the tile); they are
        // prefetched with the matrix fragments, as the panel is in stage 1
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkp - 1) * mstep;
            const int kr = min(4 * k + lk, me - 1);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(Mp, om + mk + 16u * i);
                fq[s][i] = ldo(Mq, om + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = T[kr * W + 16 * j + li];
                fy[s][j] = T[(n - 1 - kr) * W + 16 * j + li];
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkp; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        accE[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s][i], fx[s][j], accE[i][j], 0, 0, 0);
                        accO[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fq[s][i], fy[s][j], accO[i][j], 0, 0, 0);
                    }
                load(s, k0 + s + PD);
            }
        }
    }
    // stage-2 epilogue: physical rows r and n-1-r (interior row index; field row + 1)
    double* __restrict__ dst = Fq + P;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double ep[4][2];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rb0 + i) * 16 + lk + 4 * v, me - 1);
                    const int rc = half ? n - 1 - r : r;
                    ep[v][0] = nhom ? a.lift[rc + 1] : 0.0;
                    ep[v][1] = nhom ? a.lift[a.nzp + rc + 1] : 0.0;
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rb0 + i) * 16 + lk + 4 * v;
                    if (r < me && (half == 0 || r < mo)) {
                        const int row = half ? n - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            const long col = p0 + 16 * j + li;
                            const double e = accE[i][j][v], o = accO[i][j][v];
                            double x = half ? (e - o) : e + o;
                            if (nhom) x = x + ep[v][0] * vbb[j] + ep[v][1] * vbt[j];
                            if (col < P) dst[(long)row * P + col] = x;
                        }
                    }
                }
            }
        }
    }
    if (wave == 0 && lk == 0) {
        // boundary planes 0 and nz
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const long col = p0 + 16 * j + li;
            double v0 = 0.0, vn = 0.0;
            if (nhom) {
                v0 = 0.0 + a.lift[0] * vbb[j] + a.lift[a.nzp + 0] * vbt[j];
                vn = 0.0 + a.lift[a.nzp - 1] * vbb[j] + a.lift[a.nzp + a.nzp - 1] * vbt[j];
            }
            if (col < P) {
                Fq[col] = v0;
                Fq[(long)(a.nzp - 1) * P + col] = vn;
            }
        }
    }
}
// two row blocks x two column blocks per wave set (measured against 1-3 x 1-2: the
// 49 KB LDS tile per block leaves room for three blocks per CU)
constexpr int kZHRB = 2, kZHCB = 2;

// The velocity solves of the mixed form: with s = k^2 + sigma much larger than every
// |lam| (sigma = 1/(nu dt bdfv)),
//   1/(lam - s) = -1/s + c(lam),   c(lam) = lam / (s (lam - s)),   0 <= s c(lam) < 1,
// so with q = Einv r - L (L the lifting terms of the mode values) the solution is
//   u = E diag(1/(lam - s)) q + h bc = (r - E L)/(-s) + E diag(c) q + h bc.
// The first term is formed in FP64 from the right-hand side and the precomputed E L
// vectors (E Einv = I); only the correction E diag(c) q, a fraction max|lam|/s of the
// solution, goes through the two FP32 parity-split projections (v_mfma_f32_16x16x4f32,
// twice the FP64 rate), so their rounding enters the solution scaled by that fraction.
// Stage 1 leaves c q in an FP32 LDS tile; stage 2 re-reads each right-hand-side value
// just before overwriting it (the block owns its pencils; stage 1's reads are behind the
// barrier).
template <int RB, int CB, int PD, int WPE>
__global__ void __launch_bounds__(64 * kZMaxWaves) __attribute__((amdgpu_waves_per_eu(WPE)))
k_zhelm32(ZArgs a, FieldList<double> fld, FieldList<double> bct, FieldList<double> bcb) {
    extern __shared__ float Tf[];  // n rows x W pencils
    constexpr int W = 16 * CB;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int f = blockIdx.z;
    const int me = a.nout, n = a.nsp, mo = n - me;
    const int nrb = (me + 15) >> 4;
    const int base = nrb / nw, extra = nrb % nw;
    const int myn = base + (wave < extra ? 1 : 0);
    const int rb0 = wave * base + min(wave, extra);
    const long P = a.P;
    const long p0 = (long)blockIdx.x * W;
    long cj[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) cj[j] = min(p0 + 16 * j + li, P - 1);
    double* __restrict__ Fq = fld.p[f];
    double* __restrict__ X = Fq + P;  // interior rows 1..nz: read in stage 1, overwritten in stage 2
    const unsigned xstep = 4u * (unsigned)P, mstep = 4u * (unsigned)a.ldm;
    unsigned ot[CB], ob[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        ot[j] = (unsigned)lk * (unsigned)P + (unsigned)cj[j];
        ob[j] = (unsigned)(n - 1 - lk) * (unsigned)P + (unsigned)cj[j];
    }
    const unsigned om = (unsigned)(rb0 * 16 + li + lk * a.ldm);
    const int nks = (me + 3) >> 2;
    const int nkp = (nks + PD - 1) / PD * PD;
    const int kmax = (n - 4) >> 2;
    double dfac[CB], vbt[CB], vbb[CB];
    const bool nhom = bct.p[f] != nullptr;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int c = (int)cj[j];
        const double wx = a.wavx[c % a.nx], wy = a.wavy[c / a.nx];
        dfac[j] = -(wx * wx) - wy * wy - a.sigma;
        vbt[j] = nhom ? bct.p[f][c] : 0.0;
        vbb[j] = nhom ? bcb.p[f][c] : 0.0;
    }
    double rdf1[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) rdf1[j] = 1.0 / dfac[j];
    v4f accE[RB][CB], accO[RB][CB];
    float fa[PD][RB], fq[PD][RB];
    // ---------------------------------------------------------------- stage 1
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            accE[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
            accO[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    {
        const float* __restrict__ Mp = a.M32;
        const float* __restrict__ Mq = a.M32 + a.q32;
        double fx[PD][CB], fy[PD][CB];
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkp - 1) * mstep;
            const unsigned xk = (unsigned)min(k, kmax) * xstep;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(Mp, om + mk + 16u * i);
                fq[s][i] = ldo(Mq, om + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = ldo((const double*)X, ot[j] + xk);
                fy[s][j] = ldo((const double*)X, ob[j] - xk);
            }
            asm volatile("" ::: "memory");  // keep the prefetch in place (see k_zsplit)
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkp; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                float sv[CB], dv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    sv[j] = (float)(fx[s][j] + fy[s][j]);
                    dv[j] = (float)(fx[s][j] - fy[s][j]);
                }
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        accE[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[s][i], sv[j], accE[i][j], 0, 0, 0);
                        accO[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fq[s][i], dv[j], accO[i][j], 0, 0, 0);
                    }
                load(s, k0 + s + PD);
            }
        }
    }
    // stage-1 epilogue into LDS: c(lam) (Einv r - L); even mode r -> row r, odd mode r -> row n-1-r
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rb0 + i) * 16 + Z32_ROW(lk, v);
                    if (r < me && (half == 0 || r < mo)) {
                        const int row = half ? n - 1 - r : r;
                        const double lam = a.lam[row];
                        double e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0;
                        if (nhom) {
                            e1 = a.elift[row];
                            e2 = a.elift[n + row];
                            e3 = a.elift[2 * n + row];
                            e4 = a.elift[3 * n + row];
                        }
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            double x = (double)(half ? accO[i][j][v] : accE[i][j][v]);
                            if (nhom) {
                                x = x - vbb[j] * (e3 + dfac[j] * e1);
                                x = x - vbt[j] * (e4 + dfac[j] * e2);
                            }
                            // c(lam) with s = -dfac: lam / (s (lam - s)) = -(t / dfac) / (1 + t),
                            // t = lam / dfac in [0, 1/16]; it scales an FP32 value, so FP32 suffices
                            const float t = (float)(lam * rdf1[j]);
                            const float cl = -((float)rdf1[j] * t) / (1.0f + t);
                            Tf[row * W + 16 * j + li] = (float)x * cl;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // ---------------------------------------------------------------- stage 2
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            accE[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
            accO[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    {
        const float* __restrict__ Mp = a.M32b;
        const float* __restrict__ Mq = a.M32b + a.q32b;
        float fx[PD][CB], fy[PD][CB];
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkp - 1) * mstep;
            const int kr = min(4 * k + lk, me - 1);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(Mp, om + mk + 16u * i);
                fq[s][i] = ldo(Mq, om + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = Tf[kr * W + 16 * j + li];
                fy[s][j] = Tf[(n - 1 - kr) * W + 16 * j + li];
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkp; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        accE[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[s][i], fx[s][j], accE[i][j], 0, 0, 0);
                        accO[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fq[s][i], fy[s][j], accO[i][j], 0, 0, 0);
                    }
                load(s, k0 + s + PD);
            }
        }
    }
    // stage-2 epilogue: physical rows r and n-1-r; the FP64 part (r - E L)/dfac + h bc
    // plus the FP32 correction.  The right-hand-side values of one row block (both
    // halves) are loaded together before any of them is overwritten, so the loads are
    // in flight at once rather than each waiting behind the previous store.
    const double* rdf = rdf1;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < myn) {
            double rv[2][4][CB];
#pragma unroll
            for (int half = 0; half < 2; ++half)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rb0 + i) * 16 + Z32_ROW(lk, v), me - 1);
                    const int row = half ? n - 1 - r : r;
#pragma unroll
                    for (int j = 0; j < CB; ++j) rv[half][v][j] = X[(long)row * P + cj[j]];
                }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rb0 + i) * 16 + Z32_ROW(lk, v);
                    if (r < me && (half == 0 || r < mo)) {
                        const int row = half ? n - 1 - r : r;
                        double h0 = 0.0, h1 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0;
                        if (nhom) {
                            h0 = a.lift[row + 1];
                            h1 = a.lift[a.nzp + row + 1];
                            g1 = a.eel[row];
                            g2 = a.eel[n + row];
                            g3 = a.eel[2 * n + row];
                            g4 = a.eel[3 * n + row];
                        }
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            const long col = p0 + 16 * j + li;
                            const float e = accE[i][j][v], o = accO[i][j][v];
                            const double corr = (double)(half ? (e - o) : e + o);
                            double y = rv[half][v][j];
                            if (nhom) {
                                y = y - vbb[j] * (g3 + dfac[j] * g1);
                                y = y - vbt[j] * (g4 + dfac[j] * g2);
                            }
                            double x = y * rdf[j] + corr;
                            if (nhom) x = x + h0 * vbb[j] + h1 * vbt[j];
                            if (col < P) X[(long)row * P + col] = x;
                        }
                    }
                }
            }
        }
    }
    if (wave == 0 && lk == 0) {
        // boundary planes 0 and nz
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const long col = p0 + 16 * j + li;
            double v0 = 0.0, vn = 0.0;
            if (nhom) {
                v0 = 0.0 + a.lift[0] * vbb[j] + a.lift[a.nzp + 0] * vbt[j];
                vn = 0.0 + a.lift[a.nzp - 1] * vbb[j] + a.lift[a.nzp + a.nzp - 1] * vbt[j];
            }
            if (col < P) {
                Fq[col] = v0;
                Fq[(long)(a.nzp - 1) * P + col] = vn;
            }
        }
    }
}
// two row blocks x three column blocks per wave set (measured against 2-4 x 2-3 on the
// 512 x 193 x 256 step: 2 x 3 fastest; its 37 KB FP32 tile leaves room for four blocks per CU)
constexpr int kZH32RB = 2, kZH32CB = 3;

// The pressure step of navierstokes.f90:1411-1488 for a block's 16 CB pencils in
// one kernel, the intermediate field of each stage in one LDS tile (nzp rows):
//   A  r = -d(i kx u + i ky v)/dz - k^2 w          (dfdz split, div_h in the loads)
//   B  modes = Einv-split fold of r's interior rows, lifting, 1/(lam - k^2)
//   C  w_new = E-split unfold of the modes + h bc, boundary planes  (-> HBM and LDS)
//   D  y = dw_new/dz, p = -(y + div_h)/(k^2 dt bdfv), u, v -= dt bdfv (i kx, i ky) p
// Each stage repeats the operations of its stand-alone kernel (k_zsplit PRESSDIV,
// k_zhelm, k_zsplit PROJ), so the results are those of the three-kernel sequence.
struct ZPressArgs {
    const double *Dp, *Dq;                      // dfdz halves (n = nzp)
    const double *H1p, *H1q, *H2p, *H2q;        // Helmholtz halves (n = nzm)
    int ldD, ldH, nD, meD, nH, meH, xrow;       // xrow: middle row of dfdz on the VALU (-1: none)
    long P;
    int nx;
    const double* wavx;
    const double* wavy;
    const double* lam;    // parity-ordered eigenvalues, then Einv of the lifting vectors
    const double* elift;
    const double* lift;   // hbot, htop (nzp each)
    const double* bct;    // Dirichlet data of the Poisson solve (plane 0 / plane nzp-1), nullable
    const double* bcb;
    double* u;            // updated in place
    double* v;
    const double* w;      // w at the start of the pressure step
    double* wout;         // w_new (all nzp planes)
    double* p;            // pressure
    double dt, bdfv;
};

template <int RB, int CB, int PD, bool XR>
__global__ void __launch_bounds__(64 * kZMaxWaves) __attribute__((amdgpu_waves_per_eu(3)))
k_zpress(ZPressArgs a) {
    extern __shared__ double T[];  // nzp rows x W pencils
    constexpr int W = 16 * CB;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long P = a.P;
    const long p0 = (long)blockIdx.x * W;
    long cj[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) cj[j] = min(p0 + 16 * j + li, P - 1);
    // per-pencil constants (as the stand-alone kernels compute them)
    double wxc[CB], wyc[CB], sgc[CB], dfac[CB], vbt[CB], vbb[CB];
    const bool nhom = a.bct != nullptr;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int c = (int)cj[j];
        wxc[j] = a.wavx[2 * ((c % a.nx) >> 1)];
        wyc[j] = a.wavy[c / a.nx];
        sgc[j] = (c & 1) ? 1.0 : -1.0;
        const double wx = a.wavx[c % a.nx], wy = a.wavy[c / a.nx];
        dfac[j] = -(wx * wx) - wy * wy - 0.0;
        vbt[j] = nhom ? a.bct[c] : 0.0;
        vbb[j] = nhom ? a.bcb[c] : 0.0;
    }
    // row blocks of the two operator sizes
    const int nD = a.nD, meD = a.meD, moD = nD - meD;
    const int nH = a.nH, meH = a.meH, moH = nH - meH;
    const int nrbD = XR ? (meD >> 4) : ((meD + 15) >> 4);
    const int nrbH = (meH + 15) >> 4;
    const int bD = nrbD / nw, xD = nrbD % nw, mynD = bD + (wave < xD ? 1 : 0), rbD = wave * bD + min(wave, xD);
    const int bH = nrbH / nw, xH = nrbH % nw, mynH = bH + (wave < xH ? 1 : 0), rbH = wave * bH + min(wave, xH);
    const unsigned omD = (unsigned)(rbD * 16 + li + lk * a.ldD), omH = (unsigned)(rbH * 16 + li + lk * a.ldH);
    const unsigned omx = (unsigned)((XR ? a.xrow : 0) + lk * a.ldD);
    const unsigned mstD = 4u * (unsigned)a.ldD, mstH = 4u * (unsigned)a.ldH;
    const int nkD = (meD + 3) >> 2, nkpD = (nkD + PD - 1) / PD * PD, kmaxD = (nD - 4) >> 2;
    const int nkH = (meH + 3) >> 2, nkpH = (nkH + PD - 1) / PD * PD;
    const bool dox = XR && wave == 0;
    v4d accE[RB][CB], accO[RB][CB];
    double fa[PD][RB], fq[PD][RB], fx[PD][CB], fy[PD][CB], gx[PD][CB], gy[PD][CB], fxp[PD], fxq[PD];
    double xacc[CB];
    auto zero_acc = [&]() {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                accE[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
                accO[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
        for (int j = 0; j < CB; ++j) xacc[j] = 0.0;
    };
    auto mfma_step = [&](int s, const double* sv, const double* dv) {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                accE[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s][i], sv[j], accE[i][j], 0, 0, 0);
                accO[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fq[s][i], dv[j], accO[i][j], 0, 0, 0);
            }
    };
    // ------------------------------------------------------------------ stage A
    zero_acc();
    {
        const double* __restrict__ U = a.u;
        const double* __restrict__ V = a.v;
        unsigned ot[CB], ob[CB];
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const unsigned cl = (unsigned)(cj[j] ^ 1);
            ot[j] = (unsigned)lk * (unsigned)P + cl;
            ob[j] = (unsigned)(nD - 1 - lk) * (unsigned)P + cl;
        }
        const unsigned xstep = 4u * (unsigned)P;
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkpD - 1) * mstD;
            const unsigned xk = (unsigned)min(k, kmaxD) * xstep;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(a.Dp, omD + mk + 16u * i);
                fq[s][i] = ldo(a.Dq, omD + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = ldo(U, ot[j] + xk);
                fy[s][j] = ldo(U, ob[j] - xk);
                gx[s][j] = ldo(V, ot[j] + xk);
                gy[s][j] = ldo(V, ob[j] - xk);
            }
            if (XR) {
                fxp[s] = ldo(a.Dp, omx + mk);
                fxq[s] = ldo(a.Dq, omx + mk);
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkpD; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                double sv[CB], dv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    const double xt = sgc[j] * (wxc[j] * fx[s][j] + wyc[j] * gx[s][j]);
                    const double xb = sgc[j] * (wxc[j] * fy[s][j] + wyc[j] * gy[s][j]);
                    sv[j] = xt + xb;
                    dv[j] = xt - xb;
                }
                mfma_step(s, sv, dv);
                if (XR)
#pragma unroll
                    for (int j = 0; j < CB; ++j) xacc[j] = fma(fxq[s], dv[j], fma(fxp[s], sv[j], xacc[j]));
                load(s, k0 + s + PD);
            }
        }
    }
    // r = -acc + dfac w  -> LDS rows 0..nzp-1
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < mynD) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double ax[4][CB];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rbD + i) * 16 + lk + 4 * v, meD - 1);
                    const int rc = half ? nD - 1 - r : r;
#pragma unroll
                    for (int j = 0; j < CB; ++j) ax[v][j] = a.w[(long)rc * P + cj[j]];
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rbD + i) * 16 + lk + 4 * v;
                    if (r < meD && (half == 0 || r < moD)) {
                        const int row = half ? nD - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            const double e = accE[i][j][v], o = accO[i][j][v];
                            double x = half ? -1.0 * (e - o) : e + o;
                            x = -x + dfac[j] * ax[v][j];
                            T[row * W + 16 * j + li] = x;
                        }
                    }
                }
            }
        }
    }
    if (dox) {
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const double ax = a.w[(long)a.xrow * P + cj[j]];
            double x = xacc[j];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            x = -x + dfac[j] * ax;
            if (lk == 0) T[a.xrow * W + 16 * j + li] = x;
        }
    }
    __syncthreads();
    // ------------------------------------------------------------------ stage B
    // Einv-split of the interior rows 1..nz (fold around the interior mid-row)
    zero_acc();
    {
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkpH - 1) * mstH;
            const int kr = min(4 * k + lk, meH - 1);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(a.H1p, omH + mk + 16u * i);
                fq[s][i] = ldo(a.H1q, omH + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = T[(1 + kr) * W + 16 * j + li];
                fy[s][j] = T[(nH - kr) * W + 16 * j + li];
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkpH; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                double sv[CB], dv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    sv[j] = fx[s][j] + fy[s][j];
                    dv[j] = fx[s][j] - fy[s][j];
                }
                mfma_step(s, sv, dv);
                load(s, k0 + s + PD);
            }
        }
    }
    __syncthreads();  // every wave has read r
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < mynH) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double ep[4][5];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rbH + i) * 16 + lk + 4 * v, meH - 1);
                    const int rc = half ? nH - 1 - r : r;
                    ep[v][0] = a.lam[rc];
                    if (nhom) {
                        ep[v][1] = a.elift[rc];
                        ep[v][2] = a.elift[nH + rc];
                        ep[v][3] = a.elift[2 * nH + rc];
                        ep[v][4] = a.elift[3 * nH + rc];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rbH + i) * 16 + lk + 4 * v;
                    if (r < meH && (half == 0 || r < moH)) {
                        const int row = half ? nH - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            double x = half ? accO[i][j][v] : accE[i][j][v];
                            if (nhom) {
                                x = x - vbb[j] * (ep[v][3] + dfac[j] * ep[v][1]);
                                x = x - vbt[j] * (ep[v][4] + dfac[j] * ep[v][2]);
                            }
                            x = x / (ep[v][0] + dfac[j]);
                            T[row * W + 16 * j + li] = x;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // ------------------------------------------------------------------ stage C
    zero_acc();
    {
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkpH - 1) * mstH;
            const int kr = min(4 * k + lk, meH - 1);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(a.H2p, omH + mk + 16u * i);
                fq[s][i] = ldo(a.H2q, omH + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = T[kr * W + 16 * j + li];
                fy[s][j] = T[(nH - 1 - kr) * W + 16 * j + li];
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkpH; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                mfma_step(s, fx[s], fy[s]);
                load(s, k0 + s + PD);
            }
        }
    }
    __syncthreads();  // every wave has read the modes
    {
        double* __restrict__ wo = a.wout + P;  // interior rows
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            if (i < mynH) {
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    double ep[4][2];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = min((rbH + i) * 16 + lk + 4 * v, meH - 1);
                        const int rc = half ? nH - 1 - r : r;
                        ep[v][0] = nhom ? a.lift[rc + 1] : 0.0;
                        ep[v][1] = nhom ? a.lift[nD + rc + 1] : 0.0;
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = (rbH + i) * 16 + lk + 4 * v;
                        if (r < meH && (half == 0 || r < moH)) {
                            const int row = half ? nH - 1 - r : r;
#pragma unroll
                            for (int j = 0; j < CB; ++j) {
                                const long col = p0 + 16 * j + li;
                                const double e = accE[i][j][v], o = accO[i][j][v];
                                double x = half ? (e - o) : e + o;
                                if (nhom) x = x + ep[v][0] * vbb[j] + ep[v][1] * vbt[j];
                                T[(row + 1) * W + 16 * j + li] = x;
                                if (col < P) wo[(long)row * P + col] = x;
                            }
                        }
                    }
                }
            }
        }
        if (wave == 0 && lk == 0) {
            // boundary planes 0 and nz
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const long col = p0 + 16 * j + li;
                double v0 = 0.0, vn = 0.0;
                if (nhom) {
                    v0 = 0.0 + a.lift[0] * vbb[j] + a.lift[nD + 0] * vbt[j];
                    vn = 0.0 + a.lift[nD - 1] * vbb[j] + a.lift[nD + nD - 1] * vbt[j];
                }
                T[16 * j + li] = v0;
                T[(nD - 1) * W + 16 * j + li] = vn;
                if (col < P) {
                    a.wout[col] = v0;
                    a.wout[(long)(nD - 1) * P + col] = vn;
                }
            }
        }
    }
    __syncthreads();
    // ------------------------------------------------------------------ stage D
    zero_acc();
    {
        auto load = [&](int s, int k) {
            const unsigned mk = (unsigned)min(k, nkpD - 1) * mstD;
            const int kr = min(4 * k + lk, 4 * kmaxD + 3);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                fa[s][i] = ldo(a.Dp, omD + mk + 16u * i);
                fq[s][i] = ldo(a.Dq, omD + mk + 16u * i);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                fx[s][j] = T[kr * W + 16 * j + li];
                fy[s][j] = T[(nD - 1 - kr) * W + 16 * j + li];
            }
            if (XR) {
                fxp[s] = ldo(a.Dp, omx + mk);
                fxq[s] = ldo(a.Dq, omx + mk);
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
        };
#pragma unroll
        for (int s = 0; s < PD; ++s) load(s, s);
        for (int k0 = 0; k0 < nkpD; k0 += PD) {
#pragma unroll
            for (int s = 0; s < PD; ++s) {
                double sv[CB], dv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    sv[j] = fx[s][j] + fy[s][j];
                    dv[j] = fx[s][j] - fy[s][j];
                }
                mfma_step(s, sv, dv);
                if (XR)
#pragma unroll
                    for (int j = 0; j < CB; ++j) xacc[j] = fma(fxq[s], dv[j], fma(fxp[s], sv[j], xacc[j]));
                load(s, k0 + s + PD);
            }
        }
    }
    // projection at the store (lane pairs hold the Re/Im columns of a mode)
    double* __restrict__ pu = a.u;
    double* __restrict__ pv = a.v;
    double* __restrict__ pp_ = a.p;
    auto proj_store = [&](long row, int j, double y, double uc, double vc, bool st) {
        const double up = __shfl_xor(uc, 1), vp = __shfl_xor(vc, 1);
        const double divf = sgc[j] * (wxc[j] * up + wyc[j] * vp);
        const double wav2 = wxc[j] * wxc[j] + wyc[j] * wyc[j];
        double pc = 0.0;
        if (wav2 != 0.0) pc = -(y + divf) / (wav2 * a.dt * a.bdfv);
        const double pp = __shfl_xor(pc, 1);
        const double dtb = a.dt * a.bdfv;
        if (st) {
            const long o = row * P + cj[j];
            pp_[o] = pc;
            pu[o] = uc - dtb * (sgc[j] * wxc[j] * pp);
            pv[o] = vc - dtb * (sgc[j] * wyc[j] * pp);
        }
    };
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i < mynD) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double ax[4][CB], ay[4][CB];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min((rbD + i) * 16 + lk + 4 * v, meD - 1);
                    const int rc = half ? nD - 1 - r : r;
#pragma unroll
                    for (int j = 0; j < CB; ++j) {
                        ax[v][j] = pu[(long)rc * P + cj[j]];
                        ay[v][j] = pv[(long)rc * P + cj[j]];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = (rbD + i) * 16 + lk + 4 * v;
                    if (r < meD && (half == 0 || r < moD)) {
                        const int row = half ? nD - 1 - r : r;
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
                            const long col = p0 + 16 * j + li;
                            const double e = accE[i][j][v], o = accO[i][j][v];
                            const double y = half ? -1.0 * (e - o) : e + o;
                            proj_store(row, j, y, ax[v][j], ay[v][j], col < P);
                        }
                    }
                }
            }
        }
    }
    if (dox) {
        double ax[CB], ay[CB];
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            ax[j] = pu[(long)a.xrow * P + cj[j]];
            ay[j] = pv[(long)a.xrow * P + cj[j]];
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            double x = xacc[j];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            const long col = p0 + 16 * j + li;
            proj_store(a.xrow, j, x, ax[j], ay[j], lk == 0 && col < P);
        }
    }
}
constexpr int kZPRB = 2, kZPCB = 2;

// flops of one split contraction: two me x me products per column
static double zsplit_flops(const ZSplitOp& op, double cols) { return 4.0 * op.me * (double)op.me * cols; }

// ------------------------------------------------------- parity split: setup
namespace {

// column-major n x n element (i, j)
inline double cm(const std::vector<double>& m, int n, int i, int j) { return m[(size_t)j * n + i]; }

bool centro(const std::vector<double>& m, int n, double sgn, double tol) {
    double mx = 0.0;
    for (double v : m) mx = std::max(mx, std::fabs(v));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            if (std::fabs(cm(m, n, n - 1 - i, n - 1 - j) - sgn * cm(m, n, i, j)) > tol * mx) return false;
    return true;
}

// P, Q (me x me, row-major host arrays) -> device block [P | Q], column-major, ld = ldm
int zs_upload(lsb_ctx* c, ZSplitOp* op, int n, double osgn, const std::vector<double>& P,
              const std::vector<double>& Q) {
    const int me = (n + 1) / 2;
    const int ldm = ((me + 63) / 64) * 64 + 64;
    // zero columns: the k-steps (4 columns each) are padded to a multiple of the
    // prefetch depth PD; 48 = 4 * lcm(1, 2, 3, 4) covers every PD up to 4
    const int kc = (me + 47) / 48 * 48;
    std::vector<double> h((size_t)2 * ldm * kc, 0.0);
    const size_t qoff = (size_t)ldm * kc;
    for (int j = 0; j < me; ++j)
        for (int i = 0; i < me; ++i) {
            h[(size_t)j * ldm + i] = P[(size_t)i * me + j];
            h[qoff + (size_t)j * ldm + i] = Q[(size_t)i * me + j];
        }
    if (!op->d_m) LSB_TRY(dev_alloc((void**)&op->d_m, h.size() * sizeof(double)));
    LSB_HIP(hipMemcpy(op->d_m, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    std::vector<float> h32(h.begin(), h.end());
    if (!op->d_m32) LSB_TRY(dev_alloc((void**)&op->d_m32, h32.size() * sizeof(float)));
    LSB_HIP(hipMemcpy(op->d_m32, h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice));
    op->n = n;
    op->me = me;
    op->ldm = ldm;
    op->osgn = osgn;
    op->qoff = qoff;
    op->ok = true;
    return LSB_OK;
}

// Fold of a physical input for an operator M (nout = nin = n) whose rows i < me are
// used: P[i][j] = (M[i][j] + M[i][n-1-j])/2, Q[i][j] = (M[i][j] - M[i][n-1-j])/2 for
// j < mo; the middle column of odd n (s_m = 2 x_m, d_m = 0) carries M[i][m]/2 in P.
void fold_cols(const std::vector<double>& m, int n, std::vector<double>& P, std::vector<double>& Q) {
    const int me = (n + 1) / 2, mo = n / 2;
    P.assign((size_t)me * me, 0.0);
    Q.assign((size_t)me * me, 0.0);
    for (int i = 0; i < me; ++i) {
        for (int j = 0; j < mo; ++j) {
            const double a = cm(m, n, i, j), b = cm(m, n, i, n - 1 - j);
            P[(size_t)i * me + j] = 0.5 * (a + b);
            Q[(size_t)i * me + j] = 0.5 * (a - b);
        }
        if (n & 1) P[(size_t)i * me + mo] = 0.5 * cm(m, n, i, mo);
    }
}

}  // namespace

int zsplit_setup(lsb_ctx* c, const std::vector<double>& elift) {
    c->zs_dz.ok = c->zs_h1.ok = c->zs_h2.ok = false;
    const char* off = std::getenv("LSB_NO_ZSPLIT");
    if (off && off[0] && std::strcmp(off, "0") != 0) return LSB_OK;
    std::vector<double> P, Q;
    // dfdz: centro-antisymmetric on a symmetric grid
    const int nzp = c->nzp;
    if (nzp >= 8 && centro(c->mdz, nzp, -1.0, 1e-12)) {
        fold_cols(c->mdz, nzp, P, Q);
        LSB_TRY(zs_upload(c, &c->zs_dz, nzp, -1.0, P, Q));
    }
    // Helmholtz: modes of definite parity (columns of E, rows of Einv)
    const int n = c->nzm;
    if (n < 8) return LSB_OK;
    const std::vector<double>& E = c->smat;   // E[i][k]
    const std::vector<double>& Ei = c->simat; // Einv[k][j]
    std::vector<int> ev, od;
    for (int k = 0; k < n; ++k) {
        double se = 0.0, sr = 0.0, fe = 0.0, fr = 0.0, me_ = 0.0, mi = 0.0;
        for (int i = 0; i < n; ++i) {
            se += cm(E, n, i, k) * cm(E, n, n - 1 - i, k);
            sr += cm(E, n, i, k) * cm(E, n, i, k);
            fe += cm(Ei, n, k, i) * cm(Ei, n, k, n - 1 - i);
            fr += cm(Ei, n, k, i) * cm(Ei, n, k, i);
            me_ = std::max(me_, std::fabs(cm(E, n, i, k)));
            mi = std::max(mi, std::fabs(cm(Ei, n, k, i)));
        }
        if (sr <= 0.0 || fr <= 0.0) return LSB_OK;
        const double pe = se / sr, pi = fe / fr;
        const double sg = pe > 0.0 ? 1.0 : -1.0;
        if (std::fabs(pe - sg) > 1e-6 || std::fabs(pi - sg) > 1e-6) return LSB_OK;
        for (int i = 0; i < n; ++i)
            if (std::fabs(cm(E, n, n - 1 - i, k) - sg * cm(E, n, i, k)) > 1e-6 * me_ ||
                std::fabs(cm(Ei, n, k, n - 1 - i) - sg * cm(Ei, n, k, i)) > 1e-6 * mi)
                return LSB_OK;
        (sg > 0.0 ? ev : od).push_back(k);
    }
    const int me = (n + 1) / 2, mo = n / 2;
    if ((int)ev.size() != me || (int)od.size() != mo) return LSB_OK;
    // stage 1: Einv folded; even mode a -> row a (P), odd mode b -> row n-1-b (Q)
    std::vector<double> P1((size_t)me * me, 0.0), Q1((size_t)me * me, 0.0);
    for (int a = 0; a < me; ++a) {
        for (int j = 0; j < mo; ++j) P1[(size_t)a * me + j] = 0.5 * (cm(Ei, n, ev[a], j) + cm(Ei, n, ev[a], n - 1 - j));
        if (n & 1) P1[(size_t)a * me + mo] = 0.5 * cm(Ei, n, ev[a], mo);
    }
    for (int b = 0; b < mo; ++b)
        for (int j = 0; j < mo; ++j) Q1[(size_t)b * me + j] = 0.5 * (cm(Ei, n, od[b], j) - cm(Ei, n, od[b], n - 1 - j));
    // stage 2: E unfolded from parity-ordered modes
    std::vector<double> P2((size_t)me * me, 0.0), Q2((size_t)me * me, 0.0);
    for (int i = 0; i < me; ++i) {
        const bool mid = (n & 1) && i == mo;
        for (int a = 0; a < me; ++a)
            P2[(size_t)i * me + a] = mid ? cm(E, n, i, ev[a]) : 0.5 * (cm(E, n, i, ev[a]) + cm(E, n, n - 1 - i, ev[a]));
        if (!mid)
            for (int b = 0; b < mo; ++b)
                Q2[(size_t)i * me + b] = 0.5 * (cm(E, n, i, od[b]) - cm(E, n, n - 1 - i, od[b]));
    }
    // the projected pair must still invert each other
    {
        std::vector<double> Ep((size_t)n * n, 0.0), Eip((size_t)n * n, 0.0);  // row-major
        for (int i = 0; i < me; ++i) {
            for (int a = 0; a < me; ++a) {
                Ep[(size_t)i * n + ev[a]] = P2[(size_t)i * me + a];
                if (i < mo) Ep[(size_t)(n - 1 - i) * n + ev[a]] = P2[(size_t)i * me + a];
            }
            for (int b = 0; b < mo; ++b) {
                Ep[(size_t)i * n + od[b]] = Q2[(size_t)i * me + b];
                if (i < mo) Ep[(size_t)(n - 1 - i) * n + od[b]] = -Q2[(size_t)i * me + b];
            }
        }
        for (int a = 0; a < me; ++a) {
            for (int j = 0; j < mo; ++j) {
                Eip[(size_t)ev[a] * n + j] = P1[(size_t)a * me + j];
                Eip[(size_t)ev[a] * n + n - 1 - j] = P1[(size_t)a * me + j];
            }
            if (n & 1) Eip[(size_t)ev[a] * n + mo] = 2.0 * P1[(size_t)a * me + mo];
        }
        for (int b = 0; b < mo; ++b)
            for (int j = 0; j < mo; ++j) {
                Eip[(size_t)od[b] * n + j] = Q1[(size_t)b * me + j];
                Eip[(size_t)od[b] * n + n - 1 - j] = -Q1[(size_t)b * me + j];
            }
        double err = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += Ep[(size_t)i * n + k] * Eip[(size_t)k * n + j];
                err = std::max(err, std::fabs(acc - (i == j ? 1.0 : 0.0)));
            }
        if (err > 1e-9) return LSB_OK;
    }
    // eigenvalues and Einv-lifting data in parity-ordered rows, then E applied to the
    // four Einv-lifting vectors in physical rows (the FP64 part of the mixed form's solve)
    std::vector<double> hs((size_t)9 * n, 0.0);
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += cm(E, n, i, k) * elift[(size_t)q * n + k];
            hs[(size_t)5 * n + (size_t)q * n + i] = acc;
        }
    for (int a = 0; a < me; ++a) {
        hs[a] = c->diag[ev[a]];
        for (int q = 0; q < 4; ++q) hs[(size_t)n + (size_t)q * n + a] = elift[(size_t)q * n + ev[a]];
    }
    for (int b = 0; b < mo; ++b) {
        hs[n - 1 - b] = c->diag[od[b]];
        for (int q = 0; q < 4; ++q) hs[(size_t)n + (size_t)q * n + n - 1 - b] = elift[(size_t)q * n + od[b]];
    }
    if (!c->d_helm_s) LSB_TRY(dev_alloc((void**)&c->d_helm_s, hs.size() * sizeof(double)));
    LSB_HIP(hipMemcpy(c->d_helm_s, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice));
    LSB_TRY(zs_upload(c, &c->zs_h1, n, 1.0, P1, Q1));
    LSB_TRY(zs_upload(c, &c->zs_h2, n, 1.0, P2, Q2));
    return LSB_OK;
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
    const double bytes = (double)a.P * nf * (nin * sizeof(TI) + nout * sizeof(TO) * (aux ? 2 : 1));
    int tok = -1;
    if (M == c->d_mdz && zs_fits(c->zs_dz, a.P) && nout == c->nzp && nin == c->nzp) {
        int gy = 1, nw = 1;
        zsplit_shape(a, c->zs_dz, true, kZSRB32, &gy, &nw);
        dim3 grid((unsigned)((a.P + 16 * kZSCB32 - 1) / (16 * kZSCB32)), (unsigned)gy, (unsigned)nf);
        prof_mark_begin(c, PC_Z, s, &tok, zsplit_flops(c->zs_dz, (double)a.P) * nf, bytes);
        const bool xr = a.xrow >= 0;
        if (aux && xr)
            hipLaunchKernelGGL((k_zsplit_f32<TI, TO, true, true>), grid, dim3(64 * nw), 0, s, a, fi, fo, fa);
        else if (aux)
            hipLaunchKernelGGL((k_zsplit_f32<TI, TO, true, false>), grid, dim3(64 * nw), 0, s, a, fi, fo, fa);
        else if (xr)
            hipLaunchKernelGGL((k_zsplit_f32<TI, TO, false, true>), grid, dim3(64 * nw), 0, s, a, fi, fo, fa);
        else
            hipLaunchKernelGGL((k_zsplit_f32<TI, TO, false, false>), grid, dim3(64 * nw), 0, s, a, fi, fo, fa);
        prof_mark_end(c, tok, s);
        LSB_HIP(hipGetLastError());
        return LSB_OK;
    }
    a.xrow = (nout > 16 && nout % 16 == 1) ? nout - 1 : -1;
    const int nrb = (a.xrow >= 0) ? nout / 16 : (nout + 15) / 16;
    a.gblocks = std::min(nrb, kZRB32 * kZMaxWaves);
    const int gy = (nrb + a.gblocks - 1) / a.gblocks;
    const int nw = (a.gblocks + kZRB32 - 1) / kZRB32;
    dim3 grid((unsigned)((a.P + 16 * kZCB - 1) / (16 * kZCB)), (unsigned)gy, (unsigned)nf);
    prof_mark_begin(c, PC_Z, s, &tok, 2.0 * nout * nin * (double)a.P * nf, bytes);
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
template int zcontract_f32<float, double>(lsb_ctx*, const double*, int, int, int, const float* const*,
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
    const double bytes = (double)(nin + nout + (aux ? nout : 0)) * a.P * nf * sizeof(double);
    int rc;
    if (M == c->d_mdz && zs_fits(c->zs_dz, a.P) && nout == c->nzp && nin == c->nzp) {
        prof_mark_begin(c, PC_Z, s, &tok, zsplit_flops(c->zs_dz, (double)a.P) * nf, bytes);
        rc = aux ? launch_zs<ZM_AUXSUB>(a, c->zs_dz, fi, fo, fa, z, nf, s)
                 : launch_zs<ZM_PLAIN>(a, c->zs_dz, fi, fo, z, z, nf, s);
    } else {
        prof_mark_begin(c, PC_Z, s, &tok, 2.0 * nout * nin * (double)a.P * nf, bytes);
        rc = aux ? launch_z<ZM_AUXSUB>(a, fi, fo, fa, z, nf, s) : launch_z<ZM_PLAIN>(a, fi, fo, z, z, nf, s);
    }
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
    const double bytes = 3.0 * c->nzp * (double)a.P * sizeof(double);
    int rc;
    if (zs_fits(c->zs_dz, a.P)) {
        prof_mark_begin(c, PC_Z, s, &tok, zsplit_flops(c->zs_dz, (double)a.P), bytes);
        rc = launch_zs<ZM_PRESS>(a, c->zs_dz, fi, fo, aux, z, 1, s);
    } else {
        prof_mark_begin(c, PC_Z, s, &tok, 2.0 * c->nzp * (double)c->nzp * a.P, bytes);
        rc = launch_z<ZM_PRESS>(a, fi, fo, aux, z, 1, s);
    }
    prof_mark_end(c, tok, s);
    return rc;
}

bool zsplit_dz(const lsb_ctx* c) { return zs_fits(c->zs_dz, (long)c->plane_s); }

// Pressure right-hand side with the horizontal divergence formed as the contraction
// loads u and v (no div_h field): rhs = -d(i kx u + i ky v)/dz - (kx^2 + ky^2) w.
int zcontract_pressdiv(lsb_ctx* c, const double* u, const double* v, const double* w, double* rhs, hipStream_t s) {
    if (!zsplit_dz(c)) {
        set_error("zcontract_pressdiv: needs the parity-split dfdz (grid symmetric about the mid-plane)");
        return LSB_E_UNSUPPORTED;
    }
    ZArgs a{};
    a.P = (long)c->plane_s;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nzp = c->nzp;
    a.sigma = 0.0;
    a.wavx = c->d_wavx;
    a.wavy = c->d_wavy;
    FieldList<double> fi{}, fo{}, aux{}, z{};
    fi.p[0] = const_cast<double*>(u);
    fi.p[1] = const_cast<double*>(v);
    fo.p[0] = rhs;
    aux.p[0] = const_cast<double*>(w);
    int tok = -1;
    prof_mark_begin(c, PC_Z, s, &tok, zsplit_flops(c->zs_dz, (double)a.P),
                    4.0 * c->nzp * (double)a.P * sizeof(double));
    const int rc = launch_zs<ZM_PRESSDIV>(a, c->zs_dz, fi, fo, aux, z, 1, s);
    prof_mark_end(c, tok, s);
    return rc;
}

// d(phi)/dz with the projection (navierstokes.f90:1466-1488) applied as it is stored:
// p = -(dphi/dz + div_h)/(k^2 dt bdfv) (0 at k = 0), u -= dt bdfv i kx p,
// v -= dt bdfv i ky p; div_h = i kx u + i ky v from the u, v it updates.
int zcontract_proj(lsb_ctx* c, const double* phi, double* p, double* u, double* v, double dt, double bdfv,
                   hipStream_t s) {
    if (!zsplit_dz(c)) {
        set_error("zcontract_proj: needs the parity-split dfdz (grid symmetric about the mid-plane)");
        return LSB_E_UNSUPPORTED;
    }
    ZArgs a{};
    a.P = (long)c->plane_s;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nzp = c->nzp;
    a.wavx = c->d_wavx;
    a.wavy = c->d_wavy;
    a.pdt = dt;
    a.pbdfv = bdfv;
    FieldList<double> fi{}, fo{}, uv{}, z{};
    fi.p[0] = const_cast<double*>(phi);
    fo.p[0] = p;
    uv.p[0] = u;
    uv.p[1] = v;
    int tok = -1;
    prof_mark_begin(c, PC_Z, s, &tok, zsplit_flops(c->zs_dz, (double)a.P),
                    6.0 * c->nzp * (double)a.P * sizeof(double));
    const int rc = launch_zs<ZM_PROJ>(a, c->zs_dz, fi, fo, uv, z, 1, s);
    prof_mark_end(c, tok, s);
    return rc;
}

// The pressure step in one kernel (k_zpress) when the parity-split operators and
// one nzp-row LDS tile per block are available; LSB_E_UNSUPPORTED otherwise (the
// caller then runs zcontract_pressdiv, helmholtz and zcontract_proj).
int zpress_step(lsb_ctx* c, double* u, double* v, const double* w, double* wout, double* p, const double* bc_top,
                const double* bc_bot, double dt, double bdfv, hipStream_t s) {
    const long P = (long)c->plane_s;
    const ZSplitOp &D = c->zs_dz, &H1 = c->zs_h1, &H2 = c->zs_h2;
    const char* off = std::getenv("LSB_NO_ZPRESS");  // the three-kernel form, for tests
    const size_t lds = (size_t)c->nzp * 16 * kZPCB * sizeof(double);
    if (!zs_fits(D, P) || !zs_fits(H1, P) || !zs_fits(H2, P) || H1.ldm != H2.ldm || D.n != c->nzp ||
        H1.n != c->nzm || lds > 64 * 1024 || (off && off[0] && std::strcmp(off, "0") != 0))
        return LSB_E_UNSUPPORTED;
    ZPressArgs a{};
    a.Dp = D.d_m;
    a.Dq = D.d_m + D.qoff;
    a.H1p = H1.d_m;
    a.H1q = H1.d_m + H1.qoff;
    a.H2p = H2.d_m;
    a.H2q = H2.d_m + H2.qoff;
    a.ldD = D.ldm;
    a.ldH = H1.ldm;
    a.nD = D.n;
    a.meD = D.me;
    a.nH = H1.n;
    a.meH = H1.me;
    a.xrow = ((D.n & 1) && D.me > 16 && D.me % 16 == 1) ? D.me - 1 : -1;
    a.P = P;
    a.nx = c->nx;
    a.wavx = c->d_wavx;
    a.wavy = c->d_wavy;
    a.lam = c->d_helm_s;
    a.elift = c->d_helm_s + c->nzm;
    a.lift = c->d_lift;
    a.bct = bc_top;
    a.bcb = bc_bot;
    a.u = u;
    a.v = v;
    a.w = w;
    a.wout = wout;
    a.p = p;
    a.dt = dt;
    a.bdfv = bdfv;
    const int nrbD = (a.xrow >= 0) ? D.me / 16 : (D.me + 15) / 16;
    const int nrbH = (H1.me + 15) / 16;
    const int nw = (std::max(nrbD, nrbH) + kZPRB - 1) / kZPRB;
    if (nw > kZMaxWaves) return LSB_E_UNSUPPORTED;
    dim3 grid((unsigned)((P + 16 * kZPCB - 1) / (16 * kZPCB)));
    int tok = -1;
    // four split contractions per pencil; reads u, v, w, u, v; writes w_new, p, u, v
    prof_mark_begin(c, PC_HELM, s, &tok, 2.0 * (zsplit_flops(D, (double)P) + zsplit_flops(H1, (double)P)),
                    9.0 * c->nzp * (double)P * sizeof(double));
    if (a.xrow >= 0)
        hipLaunchKernelGGL((k_zpress<kZPRB, kZPCB, kZPD, true>), grid, dim3(64 * nw), lds, s, a);
    else
        hipLaunchKernelGGL((k_zpress<kZPRB, kZPCB, kZPD, false>), grid, dim3(64 * nw), lds, s, a);
    prof_mark_end(c, tok, s);
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

// In-place Helmholtz/Poisson solves of nf fields (interior rows 1..nz-1), with
// optional non-homogeneous Dirichlet data bc_top (plane 0) / bc_bot (plane nzp-1).
int helmholtz(lsb_ctx* c, double* const* rhs, const double* const* bc_top, const double* const* bc_bot, int nf,
              double sigma, hipStream_t s, bool split32) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("helmholtz: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    const int nzm = c->nzm;
    FieldList<double> fr{}, ft{}, bt{}, bb{};
    for (int f = 0; f < nf; ++f) {
        fr.p[f] = rhs[f];
        bt.p[f] = bc_top ? const_cast<double*>(bc_top[f]) : nullptr;
        bb.p[f] = bc_bot ? const_cast<double*>(bc_bot[f]) : nullptr;
    }
    // the two-kernel forms pass the mode values through a staging buffer
    auto stage_modes = [&]() -> int {
        const size_t tbytes = (size_t)nf * nzm * c->plane_s * sizeof(double);
        void* tbuf = nullptr;
        LSB_TRY(stage_get(c, 7, tbytes, &tbuf));
        for (int f = 0; f < nf; ++f) ft.p[f] = (double*)tbuf + (size_t)f * nzm * c->plane_s;
        return LSB_OK;
    };
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
    const double bytes = 2.0 * c->nzp * (double)a.P * nf * sizeof(double);
    if (zs_fits(c->zs_h1, a.P) && zs_fits(c->zs_h2, a.P) && c->zs_h1.ldm == c->zs_h2.ldm) {
        // both parity-split projections in one kernel, the mode values in LDS
        const double hflops = 2.0 * zsplit_flops(c->zs_h1, (double)a.P) * nf;
        a.lam = c->d_helm_s;
        a.elift = c->d_helm_s + nzm;
        int gy = 1, nw = 1;
        zsplit_shape(a, c->zs_h1, false, kZHRB, &gy, &nw);
        a.M3 = c->zs_h2.d_m;
        a.M4 = c->zs_h2.d_m + c->zs_h2.qoff;
        const size_t lds = (size_t)nzm * 16 * kZHCB * sizeof(double);
        dim3 grid((unsigned)((a.P + 16 * kZHCB - 1) / (16 * kZHCB)), 1, (unsigned)nf);
        const char* off = std::getenv("LSB_NO_ZHELM");  // the two-kernel form, for tests
        const bool ok = gy == 1 && lds <= 64 * 1024 && !(off && off[0] && std::strcmp(off, "0") != 0);
        const char* off32 = std::getenv("LSB_NO_ZHELM32");  // FP64 velocity solves in the mixed form, for tests
        if (ok && split32 && !(off32 && off32[0] && std::strcmp(off32, "0") != 0)) {
            // the FP32-correction form needs sigma to dominate every eigenvalue: then
            // s c(lam) <= max|lam| / sigma <= 1/16 scales the FP32 rounding
            double lmax = 0.0;
            for (int k = 0; k < nzm; ++k) lmax = std::max(lmax, std::fabs(c->diag[k]));
            int gy32 = 1, nw32 = 1;
            ZArgs b = a;
            zsplit_shape(b, c->zs_h1, false, kZH32RB, &gy32, &nw32);
            b.M32b = c->zs_h2.d_m32;
            b.q32b = (long)c->zs_h2.qoff;
            b.eel = c->d_helm_s + 5 * nzm;
            const size_t lds32 = (size_t)nzm * 16 * kZH32CB * sizeof(float);
            if (sigma >= 16.0 * lmax && gy32 == 1 && lds32 <= 64 * 1024) {
                // profiled with the FP32 contractions (its projections run on the FP32 matrix cores)
                prof_mark_begin(c, PC_Z, s, &tok, hflops, bytes);
                dim3 grid32((unsigned)((a.P + 16 * kZH32CB - 1) / (16 * kZH32CB)), 1, (unsigned)nf);
                hipLaunchKernelGGL((k_zhelm32<kZH32RB, kZH32CB, 3, 2>), grid32, dim3(64 * nw32), lds32, s, b, fr,
                                   bt, bb);
                prof_mark_end(c, tok, s);
                LSB_HIP(hipGetLastError());
                return LSB_OK;
            }
        }
        prof_mark_begin(c, PC_HELM, s, &tok, hflops, bytes);
        if (ok) hipLaunchKernelGGL((k_zhelm<kZHRB, kZHCB, 3, 3>), grid, dim3(64 * nw), lds, s, a, fr, bt, bb);
        if (ok) {
            prof_mark_end(c, tok, s);
            LSB_HIP(hipGetLastError());
            return LSB_OK;
        }
        // larger nz: two kernels through the mode-space buffer
        LSB_TRY(stage_modes());
        a.lam = c->d_helm_s;
        a.elift = c->d_helm_s + nzm;
        a.in_row0 = 1;
        a.out_row0 = 0;
        int rc = launch_zs<ZM_HELM1>(a, c->zs_h1, fr, ft, bt, bb, nf, s);
        if (rc == LSB_OK) {
            a.in_row0 = 0;
            a.out_row0 = 1;
            rc = launch_zs<ZM_HELM2>(a, c->zs_h2, ft, fr, bt, bb, nf, s);
        }
        prof_mark_end(c, tok, s);
        return rc;
    }
    // two nzm x nzm projections per pencil; read rhs and write the solution once each
    LSB_TRY(stage_modes());
    prof_mark_begin(c, PC_HELM, s, &tok, 4.0 * nzm * (double)nzm * a.P * nf, bytes);
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
