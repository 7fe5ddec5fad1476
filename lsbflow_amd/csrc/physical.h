// Argument blocks and launchers of the physical-space kernels.
#pragma once

#include "lsb_internal.h"

namespace lsb {

enum { NL_NONE = 0, NL_SIGMA = 1, NL_STATIC = 2, NL_EXTTAU = 3, NL_PREP = 4 };

// T is the storage type of the physical fields (double; float in the mixed form).
template <class T>
struct NlArgsT {
    const T* ui[3];        // velocity perturbation (physical)
    const T* g[9];         // du_l/dx_m at l + 3m (physical)
    const double* ubar;    // nzp
    const double* ubarp;   // nzp
    const double* delta;   // nzp (sigma model filter width)
    const double* smag;    // nzp (static model factor)
    const T* tau[9];       // NL_EXTTAU: precomputed tau (nullable entries = 0)
    T* nlt[3];             // -1/2 u_j du_i/dx_j
    T* h[9];               // h_lm = 1/2 u_l u_m + tau_lm (NL_PREP: u_l u_m); null entries skipped
    T* uo[3];              // NL_PREP: velocity with ubar added (may alias ui)
    T* go[9];              // NL_PREP: g with ubarp added at (0, 2) (may alias g)
    const double* lam;     // sponge coefficient lambda(x), nx values, or null (forcing = 0)
    int write_base;
    int sub_base;          // store h_00 - ubar^2/2 (the plane-constant base part, restored in spectral space)
};
typedef NlArgsT<double> NlArgs;

struct DynArgs {
    const double* g[9];
    const double* delta;   // nzp (dynamic model widths)
    double* snorm;
    double* sij[9];
    double* l1[9];
    const double* sijhat[9];
    const double* l1hat[9];
    const double* uihat[3];
    const double* uuhat[9];
    double* num;
    double* den;
    const double* numf;
    const double* denf;
    double* tau[9];
};

int ph_zero_pads(lsb_ctx* c, double* const* f, int nf, hipStream_t s);
// qb = (first ? 0 : qb) + coef*q on the physical points of nf fields
int ph_axpy(lsb_ctx* c, const double* const* q, double* const* qb, int nf, double coef, bool first, hipStream_t s);
// e_host[k] = sum over the nf fields and the physical points of plane k of (q - f)^2 (synchronises)
int ph_plane_energy(lsb_ctx* c, const double* const* q, const double* const* f, int nf, double* e_host,
                    hipStream_t s);
int ph_nonlinear(lsb_ctx* c, int mode, const NlArgs& a, hipStream_t s);
int ph_nonlinear_f32(lsb_ctx* c, int mode, const NlArgsT<float>& a, hipStream_t s);
int ph_hcombine(lsb_ctx* c, const double* const* uu, const double* const* tau, double* const* h, int n,
                hipStream_t s);
int ph_smag_factor(lsb_ctx* c, const double* g02, const double* ubarp, double cs, hipStream_t s);
int ph_smag_factor_f32(lsb_ctx* c, const float* g02, const double* ubarp, double cs, hipStream_t s);
int ph_dyn(lsb_ctx* c, int stage, const DynArgs& a, hipStream_t s);
int ph_interscale(lsb_ctx* c, int stage, const FieldList<double>& p0, const FieldList<double>& p1,
                  const FieldList<double>& p2, const FieldList<double>& p3, const FieldList<double>& p4,
                  hipStream_t s);

}  // namespace lsb
