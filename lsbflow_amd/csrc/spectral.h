// Argument blocks and launchers of the fused spectral step kernels.
#pragma once

#include "lsb_internal.h"

namespace lsb {

struct StepRhsArgs {
    // the four pieces of this step's N (double, or float when n32 = 1: mixed form)
    const void* nlt[3];    // forward FFT of the convective term
    const void* h0[3];     // hij(l, x) spectral
    const void* h1[3];     // hij(l, y) spectral
    const void* hz[3];     // dfdz(hij(l, z)); null when already folded into nlt
    int n32;
    // N history (nout, nl, nlm): double, or float when h32 = 1 (mixed form: N is
    // FP32-accurate there, so its history is kept in FP32)
    int h32;
    void* nout[3];         // N_l (becomes nlnif)
    const void* nl[3];     // nlnif
    const void* nlm[3];    // nlnmif
    const double* u[3];
    const double* un[3];
    const double* unm[3];
    double* rhs[3];
    const double* ubarpp;  // nzp (device)
    double ab0, ab1, ab2, bdf0, bdf1, bdf2, dt, dtnui;
};

struct StepBcArgs {
    const double* rhs0;    // plane 0 of rhs_u
    const double* rhs1;    // plane 0 of rhs_v
    const double* p_wall;  // plane nzp-1 of pf
    const double* pn_wall; // plane nzp-1 of pnf
    const double* p_top;   // plane 0 of pf  (top_bc = 1)
    const double* pn_top;  // plane 0 of pnf (top_bc = 1)
    const double* d2w;     // 2 planes: rows 0 and nzp-1 of d2fdz2(w^n)
    const double* bnd_top; // bound() spectral plane 0 (nullable)
    const double* bnd_bot; // bound() spectral plane nzp-1 (nullable)
    double* bct[3];
    double* bcb[3];
    double dtnui, bdfv, dt, xnu, pfac0, pfac1;
    int nhom;
    int top_bc;
    double* cdiff[2];      // top_bc 2: Dirichlet minus free ceiling values of u, v
};

// top_bc 2: bct[l] += F[m(x) F^-1(cdiff[l])] for l = 0, 1 (physical planes as scratch)
int sp_ceiling_blend(lsb_ctx* c, double* const* bct, double* const* cdiff, double* const* phys, const double* m,
                     hipStream_t s);

int sp_rhs_vel(lsb_ctx* c, const StepRhsArgs& a, hipStream_t s);
int sp_bc_vel(lsb_ctx* c, const StepBcArgs& a, hipStream_t s);
int sp_project(lsb_ctx* c, const double* dwdz, const double* divf, double* p, double* u, double* v, double dt,
               double bdfv, hipStream_t s);
const double* sp_filter_table(lsb_ctx* c, double kc, double p, double amp, int filtery, hipStream_t s);
int sp_specfilter(lsb_ctx* c, const double* const* in, double* const* out, int nf, double kc, double p, double amp,
                  int filtery, hipStream_t s);

}  // namespace lsb
