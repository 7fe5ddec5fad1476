// Argument blocks and launchers of the physical-space kernels.
#pragma once

#include <cmath>

#include "lsb_internal.h"

namespace lsb {

// x / s of the 3-point filter passes.  When s = 2a + b is a power of two, x * (1/s) is the
// same correctly rounded value (1/s is exact), without the FP64 division sequence.
__device__ __forceinline__ double fdiv(double x, double s, double inv, bool pow2) { return pow2 ? x * inv : x / s; }
__host__ __device__ inline bool pow2_exact(double s) {
    int e = 0;
    return s > 0.0 && std::frexp(s, &e) == 0.5;
}

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
    const double* cplane;  // plane means <M:L>(k), <M:M>(k) (2 nzp) instead of numf/denf, or null
    double* tau[9];
    double* h[9];          // non-null: h = uu/2 + tau written over uu (the distinct components)
    const double* uu[9];
    int hsym;              // h/uu are the symmetric view (6 distinct fields)
};

// mixed-form dynamic model (dynamic.hip): FP32 fields, symmetric tensors as 6 fields
// in the order (0,0), (1,0), (2,0), (1,1), (2,1), (2,2)
struct DynMixArgs {
    const float* ui[3];    // perturbation velocity
    const float* g[9];     // perturbation gradient at l + 3m
    const double* ubar;    // nzp
    const double* ubarp;   // nzp
    const double* delta;   // nzp (dynamic widths)
    const double* lam;     // sponge lambda(x) or null
    float* nlt[3];         // stage 1 outputs
    float* p[6];
    float* sp[6];
    float* l1[6];
    float* snorm;
    const float* sph[6];   // stage 2 inputs (filtered)
    const float* l1h[6];
    const float* uh[3];
    const float* ph[6];
    const double* base;    // 3 nzp: U^, (S^U)^_02, (U^2)^ - (U^)^2
    float* num;
    float* den;
    const float* numf;     // stage 3 inputs
    const float* denf;
    const double* cplane;  // or the plane means (2 nzp)
    float* h[6];           // holds p on entry, h = p/2 + tau on exit
};

int ph_zero_pads(lsb_ctx* c, double* const* f, int nf, hipStream_t s);
int ph_dyn_mixed(lsb_ctx* c, int stage, const DynMixArgs& a, hipStream_t s);
// the (1, 1) average of num and den and stage 3 in one pass (numf, denf not written);
// LSB_E_UNSUPPORTED where that pass does not apply (the caller runs the two steps)
int ph_dyn_mixed_avg3(lsb_ctx* c, const DynMixArgs& a, hipStream_t s);
// out[k] = mean of f0 over plane k, out[nzp + k] = mean of f1 (nx physical points per row);
// scratch: ph_plane_mean2_scratch(c) doubles
template <class T>
int ph_plane_mean2(lsb_ctx* c, const T* f0, const T* f1, double* scratch, double* out, hipStream_t s);
size_t ph_plane_mean2_scratch(const lsb_ctx* c);
void dyn_base_profiles(const lsb_ctx* c, std::vector<double>& out);
int ph_filter3_f32(lsb_ctx* c, const float* const* in, float* const* out, int nf, double a, double b, hipStream_t s);
// time the fused filter's launch variants (FP64 and FP32) on nf scratch fields and record them
int ph_tune_filt3(lsb_ctx* c, void* work, int nf, hipStream_t s);
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
