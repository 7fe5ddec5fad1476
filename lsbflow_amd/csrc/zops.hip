// Wall-normal (Chebyshev) operators as batched dense contractions.
//
//   dfdz / d2fdz2 / dpdz  (derivatives.f90:112-188 -> chebyshev.f90:204-270):
//       out[k][p] = sum_m M[k][m] in[m][p]
//   solve_helmholtz       (helmholtz.f90:130-208), two contractions:
//       t[r][p]   = (sum_m Einv[r][m] q[m+1][p]) / (lam_r - kx^2 - ky^2 - sigma)
//       q[r+1][p] = sum_m E[r][m] t[m][p],   q[0][p] = q[nz][p] = 0
//   solve_nhom_helmholtz  (helmholtz.f90:211-358): the lifting subtraction is
//       fused into the stage-1 load and the h*bc addition into the stage-2 store.
//
// A spectral field is a row-major [nzp][P] matrix with P = nx*ny reals, so every
// contraction is M (nout x nin) times a tall skinny panel; each (Re or Im) column
// is one (kx,ky) pencil.  FP64 path: one wave per ROWS output rows and 64
// columns, the matrix column M[:, m] read through the scalar cache (wave-uniform
// address) and the panel element broadcast into ROWS FMAs.  On gfx950 the FP64
// vector FMA rate equals the FP64 MFMA rate (78.6 TFLOP/s), so this form runs at
// the same roofline as v_mfma_f64 without the operand shuffles.

#include "lsb_internal.h"

namespace lsb {

enum { ZM_PLAIN = 0, ZM_HELM1 = 1, ZM_HELM2 = 2 };

struct ZArgs {
    const double* M;    // column-major, leading dim ldm (zero padded rows)
    int ldm, nout, nin;
    int in_row0;        // first input plane
    int out_row0;       // first output plane
    long P;             // columns per plane (nx*ny)
    int nx, ny;
    int nwaves;
    double sigma;
    const double* wavx;   // nx
    const double* wavy;   // ny
    const double* lam;    // nzm (stage 1)
    const double* lift;   // hbot, htop, d2hbot, d2htop (4*nzp)
    int nzp;
};

template <int ROWS, int MODE>
__global__ void __launch_bounds__(1024) k_zcontract(ZArgs a, FieldList<double> in, FieldList<double> out,
                                                    FieldList<double> bct, FieldList<double> bcb) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const long col = (long)blockIdx.x * 64 + lane;
    const int f = blockIdx.z;
    const int r0 = (blockIdx.y * a.nwaves + wave) * ROWS;
    if (r0 >= a.nout) return;
    const bool valid = col < a.P;
    const long cc = valid ? col : 0;
    const double* __restrict__ src = in.p[f] + (long)a.in_row0 * a.P;
    double dfac = 0.0;
    double vbt = 0.0, vbb = 0.0;
    const bool nhom = (MODE != ZM_PLAIN) && (bct.p[f] != nullptr);
    if (MODE != ZM_PLAIN) {
        const int i = (int)(cc % a.nx);
        const int j = (int)((cc / a.nx) % a.ny);
        const double wx = a.wavx[i], wy = a.wavy[j];
        dfac = -(wx * wx) - wy * wy - a.sigma;
        if (nhom) {
            vbt = bct.p[f][cc];
            vbb = bcb.p[f][cc];
        }
    }
    double acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = 0.0;
    const double* __restrict__ M = a.M + r0;
    for (int m = 0; m < a.nin; ++m) {
        double x = src[(long)m * a.P + cc];
        if (MODE == ZM_HELM1 && nhom) {
            const int k = m + 1;  // physical plane of interior row m
            const double abot = a.lift[2 * a.nzp + k] + dfac * a.lift[k];
            x = x - abot * vbb;
            const double atop = a.lift[3 * a.nzp + k] + dfac * a.lift[a.nzp + k];
            x = x - atop * vbt;
        }
        const double* __restrict__ Mc = M + (long)m * a.ldm;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) acc[r] = fma(Mc[r], x, acc[r]);
    }
    if (!valid) return;
    double* __restrict__ dst = out.p[f] + (long)a.out_row0 * a.P;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int row = r0 + r;
        if (row < a.nout) {
            double v = acc[r];
            if (MODE == ZM_HELM1) v = v / (a.lam[row] + dfac);
            if (MODE == ZM_HELM2 && nhom) {
                const int k = row + 1;
                v = v + a.lift[k] * vbb + a.lift[a.nzp + k] * vbt;
            }
            dst[(long)row * a.P + col] = v;
        }
    }
    if (MODE == ZM_HELM2 && r0 == 0) {
        // boundary planes 0 and nz of the full field (out_row0 == 1)
        double* full = out.p[f];
        double v0 = 0.0, vn = 0.0;
        if (nhom) {
            v0 = 0.0 + a.lift[0] * vbb + a.lift[a.nzp + 0] * vbt;
            vn = 0.0 + a.lift[a.nzp - 1] * vbb + a.lift[a.nzp + a.nzp - 1] * vbt;
        }
        full[col] = v0;
        full[(long)(a.nzp - 1) * a.P + col] = vn;
    }
}

static void pick_rows(int nout, int* rows, int* nwaves) {
    // row blocks above 48 exceed the 128 VGPRs a 1024-thread block allows and spill
    static const int cand[5] = {48, 40, 32, 24, 16};
    int best = -1, bw = 0, bwaste = 1 << 30;
    for (int ci = 0; ci < 5; ++ci) {
        int r = cand[ci];
        int nw = (nout + r - 1) / r;
        if (nw > 16) continue;
        int waste = nw * r - nout;
        if (waste < bwaste) {
            bwaste = waste;
            best = r;
            bw = nw;
        }
    }
    if (best < 0) {
        best = 48;
        bw = 16;
    }
    *rows = best;
    *nwaves = bw;
}

template <int MODE>
static int launch_z(const ZArgs& a0, const FieldList<double>& in, const FieldList<double>& out,
                    const FieldList<double>& bt, const FieldList<double>& bb, int nf, hipStream_t s) {
    ZArgs a = a0;
    int rows, nw;
    pick_rows(a.nout, &rows, &nw);
    const int gy = (a.nout + rows * nw - 1) / (rows * nw);
    a.nwaves = nw;
    dim3 grid((unsigned)((a.P + 63) / 64), (unsigned)gy, (unsigned)nf);
    dim3 block(64 * nw);
#define LSB_ZL(R)                                                                                  \
    case R:                                                                                        \
        hipLaunchKernelGGL((k_zcontract<R, MODE>), grid, block, 0, s, a, in, out, bt, bb); \
        break;
    switch (rows) {
        LSB_ZL(48)
        LSB_ZL(40)
        LSB_ZL(32)
        LSB_ZL(24)
        LSB_ZL(16)
        default:
            set_error("zcontract: bad row block");
            return LSB_E_ARG;
    }
#undef LSB_ZL
    LSB_HIP(hipGetLastError());
    return LSB_OK;
}

int zcontract(lsb_ctx* c, const double* M, int ldm, int nout, int nin, const double* const* in,
              double* const* out, int nf, hipStream_t s) {
    if (nf < 1 || nf > kMaxFields) {
        set_error("zcontract: 1..16 fields per launch");
        return LSB_E_ARG;
    }
    ZArgs a{};
    a.M = M;
    a.ldm = ldm;
    a.nout = nout;
    a.nin = nin;
    a.P = (long)c->plane_s;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nzp = c->nzp;
    FieldList<double> fi{}, fo{}, z{};
    for (int f = 0; f < nf; ++f) {
        fi.p[f] = const_cast<double*>(in[f]);
        fo.p[f] = out[f];
        z.p[f] = nullptr;
    }
    int tok = -1;
    prof_mark_begin(c, PC_Z, s, &tok, 2.0 * nout * nin * (double)a.P * nf,
                    (double)(nin + nout) * a.P * nf * sizeof(double));
    int rc = launch_z<ZM_PLAIN>(a, fi, fo, z, z, nf, s);
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
